// Shared host/device definitions of libmfl (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/mfl.h"

namespace mfl {

constexpr uint32_t kBuckets = 262144u;         // libCL/phys/oclFluid3D.cl:32, oclFluid3D.cpp:17
constexpr uint32_t kEmpty = 0xFFFFFFFFu;       // libCL/phys/oclFluid3D.cpp:309
constexpr uint32_t kReferenceClear = 65536u;   // 262,144-byte fill of uint entries, cpp:310-312
constexpr uint32_t kGroup = 256u;              // cLocalSize, cpp:16
constexpr int kWave = 64;
constexpr uint32_t kXcds = 8;                  // XCDs of an MI355X, each with its own L2

// ---- error reporting (mfl_last_error) ---------------------------------------------------------
void set_error(const std::string& msg);
bool hip_ok(hipError_t e, const char* what);

#define MFL_HIP(call)                         \
    do {                                      \
        if (!::mfl::hip_ok((call), #call)) return 0; \
    } while (0)

// ---- device memory owned by a program ----------------------------------------------------------
// Move-only: a program's buffers are freed with the program (caller-owned memory is never freed).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool owned = true;

    DeviceBuffer() = default;
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes), owned(o.owned)
    {
        o.ptr = nullptr;
        o.bytes = 0;
        o.owned = true;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o) {
            release();
            ptr = o.ptr;
            bytes = o.bytes;
            owned = o.owned;
            o.ptr = nullptr;
            o.bytes = 0;
            o.owned = true;
        }
        return *this;
    }

    bool allocate(size_t nbytes);   // (re)allocates when owned; contents lost
    bool fit(size_t nbytes);        // grow-only (oclRadixSort::fit, libCL/sort/oclRadixSort.cpp:50-59)
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// ---- cell hash (libCL/phys/oclFluid3D.cl:34-52) ------------------------------------------------
// (int)floor(p / h) with the IEEE divide (HIP's default fp32 divide is correctly rounded), and the
// uint-wrapping XOR of three prime products reduced mod 2^18.
// Tile of this block when consecutive tiles should share an L2: the hardware deals a grid's blocks to
// the 8 XCDs round-robin (block b runs on XCD b % 8), and the tiles are dealt so that every XCD takes
// one contiguous range of them.
__device__ inline uint32_t xcd_tile()
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// Exclusive scan of one value per thread over a block of THREADS threads (a multiple of 64; s_wave holds
// THREADS / 64 words of LDS); *total receives the block's sum. Every thread of the block must call it.
template <int THREADS>
__device__ inline uint32_t block_scan(uint32_t v, uint32_t* s_wave, uint32_t* total)
{
    constexpr int kWaves = THREADS / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    if (lane == kWave - 1) s_wave[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        const uint32_t s = s_wave[w];
        before += (w < wave) ? s : 0u;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// Appends the bytes of v to a launch signature (the state a captured graph depends on): comparing
// raw bytes costs far less per step than formatting the values.
template <class T>
inline void sig_add(std::string& sig, const T& v)
{
    sig.append(reinterpret_cast<const char*>(&v), sizeof(T));
}

__host__ __device__ inline int grid_coord(float p, float h) { return (int)floorf(p / h); }

__host__ __device__ inline uint32_t cell_hash(int gx, int gy, int gz)
{
    const uint32_t n = (73856093u * (uint32_t)gx) ^ (19349663u * (uint32_t)gy) ^ (83492791u * (uint32_t)gz);
    return n & (kBuckets - 1u);
}

// The least float T with sqrtf(T) >= h. IEEE square root is correctly rounded and non-decreasing, so
// for every r2 >= 0: sqrtf(r2) < h exactly when r2 < T, the reference's fast_length(d) < h test in the
// exact form the library defines it, as one comparison (host and device sqrt round alike).
inline float sqrt_cut(float h)
{
    if (!(h > 0.f)) return h > 0.f ? h : (h == h ? 0.f : h);  // h <= 0: never accepted; NaN: never
    float x = h * h;
    while (x > 0.f && std::sqrt(x) >= h) x = std::nextafter(x, 0.f);
    while (std::sqrt(x) < h) x = std::nextafter(x, __builtin_inff());
    return x;
}

// |d|^2 as fma(dz, dz, fma(dy, dy, dx * dx)). Every kernel forms squared distances with it (the
// library is built with contraction off), so acceptance decisions and sums agree bit for bit between
// the kernels that share a pass.
__host__ __device__ inline float dist2(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }

}  // namespace mfl
