// Stable LSD key/value radix sort, gfx950 / wave64.
//
// Same contract as oclRadixSort::compute (libCL/sort/oclRadixSort.cpp:62-144): after the call the
// pairs are stably ordered by bits [start_bit, end_bit) of the key. The reference walks 4-bit digits
// with a serial 256-long scan per 256-item block (libCL/sort/oclRadixSort.cl:20-222); here each pass
// takes up to 9 bits (an 18-bit key sorts in two passes) and has three kernels:
//   k_tile_counts : per-tile digit histogram in LDS            (reads 4 B/key)
//   k_scan_digits : per-digit exclusive scan over tiles          (tiny)
//   k_scatter     : stable in-tile ranking with wave64 ballot matching (no serial scans), tile
//                   reordered by digit in LDS, then runs written contiguously (reads 8, writes 8 B/key)
#include "radix_sort.hpp"
#include "tile_rank.hpp"

#include <mutex>

namespace mfl {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kMaxDigitBits = 9;
constexpr uint32_t kMaxRadix = 1u << kMaxDigitBits;  // 512 digits: two per thread in the digit loops

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t* total)
{
    return block_scan<kThreads>(v, s_wave, total);
}

// Tiles are dealt to the XCDs in contiguous ranges (xcd_tile): neighbouring tiles write neighbouring
// words of the count table and adjoining pieces of each digit's run, so their partial cache lines meet
// in one L2 instead of being written back from eight.
template <int ITEMS>
__global__ __launch_bounds__(kThreads) void k_tile_counts(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                          int bits, uint32_t* __restrict__ counts, uint32_t ntiles)
{
    __shared__ uint32_t hist[kMaxRadix];
    const uint32_t radix = 1u << bits, mask = radix - 1u;
    for (uint32_t d = threadIdx.x; d < radix; d += kThreads) hist[d] = 0;
    __syncthreads();
    const uint32_t tile = xcd_tile();
    const size_t base = (size_t)tile * (kThreads * ITEMS);
    if (base + kThreads * ITEMS <= n && (reinterpret_cast<uintptr_t>(keys) & 15u) == 0) {
        // a whole tile of a 16-byte aligned array, read 16 bytes per load (a histogram does not depend
        // on the order its keys are counted in)
        // every load issued before the first count (else each load waits for the previous one's counts)
        const uint4* k4 = reinterpret_cast<const uint4*>(keys + base);
        uint4 q[ITEMS / 4];
#pragma unroll
        for (int k = 0; k < ITEMS / 4; k++) q[k] = k4[k * kThreads + threadIdx.x];
#pragma unroll
        for (int k = 0; k < ITEMS / 4; k++) {
            atomicAdd(&hist[(q[k].x >> shift) & mask], 1u);
            atomicAdd(&hist[(q[k].y >> shift) & mask], 1u);
            atomicAdd(&hist[(q[k].z >> shift) & mask], 1u);
            atomicAdd(&hist[(q[k].w >> shift) & mask], 1u);
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const size_t i = base + (size_t)k * kThreads + threadIdx.x;
            if (i < n) atomicAdd(&hist[(keys[i] >> shift) & mask], 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < radix; d += kThreads) counts[(size_t)d * ntiles + tile] = hist[d];
}

// One block per digit: row d of the digit-major [radix x ntiles] table becomes its exclusive scan.
__global__ __launch_bounds__(kThreads) void k_scan_digits(uint32_t* __restrict__ counts, uint32_t ntiles,
                                                          uint32_t* __restrict__ totals)
{
    // rounds of kThreads x 4 consecutive entries (thread t takes 4 t .. 4 t + 3 of a round: 16-byte
    // strides across the lanes); the loads of four rounds are in flight together (clamped indices)
    constexpr uint32_t ROUND = kThreads * 4, ROUNDS = 4;
    __shared__ uint32_t s_wave[kWaves];
    uint32_t* row = counts + (size_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += ROUND * ROUNDS) {
        uint32_t c[ROUNDS][4];
#pragma unroll
        for (uint32_t r = 0; r < ROUNDS; r++)
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) c[r][k] = row[min(base + r * ROUND + threadIdx.x * 4 + k, ntiles - 1)];
#pragma unroll
        for (uint32_t r = 0; r < ROUNDS; r++) {
            const uint32_t i0 = base + r * ROUND + threadIdx.x * 4;
            if (base + r * ROUND >= ntiles) break;  // block-uniform
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) {
                if (i0 + k >= ntiles) c[r][k] = 0u;
                sum += c[r][k];
            }
            uint32_t tot;
            uint32_t run = carry + block_exclusive_scan(sum, s_wave, &tot);
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) {
                if (i0 + k < ntiles) row[i0 + k] = run;
                run += c[r][k];
            }
            carry += tot;
        }
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// ---- one tile of a pass: keys of wave w, round r, lane l sit at tile offset (w*ITEMS + r)*64 + l, so
// (wave, round, lane) order is input order; a key's rank in the tile is
//   [same digit in earlier waves] + [same digit in earlier rounds of its wave] + [lower lanes, same round].

template <int ITEMS, bool IDENTITY_VALS>
__device__ __forceinline__ void load_tile(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t n,
                                 size_t tile_base, uint32_t first, uint32_t* key, uint32_t* val)
{
    if (tile_base + kThreads * ITEMS <= n) {  // a whole tile: no bound tests
        const uint32_t* kt = kin + tile_base + first;
        const uint32_t* vt = vin + tile_base + first;
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            // read once: non-temporal, so that the partial lines of the tile's output runs stay in L2
            key[r] = __builtin_nontemporal_load(kt + r * kWave);
            val[r] = IDENTITY_VALS ? (uint32_t)(tile_base + first + r * kWave) : __builtin_nontemporal_load(vt + r * kWave);
        }
    } else {
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const size_t idx = tile_base + first + (size_t)r * kWave;
            const bool valid = idx < n;
            key[r] = valid ? kin[idx] : 0u;
            if (IDENTITY_VALS)
                val[r] = (uint32_t)idx;
            else
                val[r] = valid ? vin[idx] : 0u;
        }
    }
}

// The tile reordered by digit in LDS, then written out in runs: the key at tile position p, digit d,
// goes to s_base[d] + (p - s_start[d]), s_start[d] being digit d's first position in the tile.
template <int ITEMS, uint32_t R>
__device__ __forceinline__ void write_runs(const uint32_t* key, const uint32_t* val, const uint32_t* rank, uint32_t n,
                                  size_t tile_base, int shift, uint32_t mask, const uint32_t* s_cnt,
                                  const uint32_t* s_start, const uint32_t* s_base, uint32_t* s_keys,
                                  uint32_t* s_vals, uint32_t* __restrict__ kout, uint32_t* __restrict__ vout)
{
    constexpr int TILE = kThreads * ITEMS;
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wave = t / kWave;
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const size_t idx = tile_base + ((size_t)(wave * ITEMS + r) << 6) + lane;
        if (idx < n) {
            const uint32_t d = (key[r] >> shift) & mask;
            const uint32_t pos = s_start[d] + s_cnt[wave * R + d] + rank[r];
            s_keys[pos] = key[r];
            s_vals[pos] = val[r];
        }
    }
    __syncthreads();
    const size_t remain = n - tile_base;
    const uint32_t tile_n = remain < (size_t)TILE ? (uint32_t)remain : (uint32_t)TILE;
    for (uint32_t p = t; p < tile_n; p += kThreads) {
        const uint32_t k = s_keys[p];
        const uint32_t d = (k >> shift) & mask;
        const uint32_t o = s_base[d] + (p - s_start[d]);
        if (o < n) {  // always true for a consistent count table; keeps a corrupt one in bounds
            kout[o] = k;
            vout[o] = s_vals[p];
        }
    }
}

// Stable scatter of one tile by the scanned count table. BITS > 0 fixes the digit width at compile time
// (the common 7, 8 and 9); BITS = 0 takes `bits_rt`.
template <int ITEMS, bool IDENTITY_VALS, int BITS>
__global__ __launch_bounds__(kThreads) void k_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                      uint32_t n, int shift, int bits_rt,
                                                      const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ totals, uint32_t ntiles)
{
    const int bits = BITS > 0 ? BITS : bits_rt;
    constexpr int TILE = kThreads * ITEMS;
    constexpr uint32_t R = BITS > 0 ? 1u << BITS : kMaxRadix;  // digit table size (LDS per block)
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_cnt[kWaves * R];
    __shared__ uint32_t s_start[R];
    __shared__ uint32_t s_base[R];
    __shared__ uint32_t s_wave[kWaves];

    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wave = t / kWave;
    const uint32_t radix = 1u << bits, mask = radix - 1u;
    const uint32_t tile = xcd_tile();
    const size_t tile_base = (size_t)tile * TILE;

    // the tile's loads first: they are in flight while the digit bases are read and scanned
    uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
    const uint32_t first = (uint32_t)(wave * ITEMS) * kWave + (uint32_t)lane;  // tile offset of item 0
    load_tile<ITEMS, IDENTITY_VALS>(kin, vin, n, tile_base, first, key, val);

    // thread t owns digits 2t and 2t + 1 in the per-digit loops
    const uint32_t d0 = 2u * (uint32_t)t, d1 = d0 + 1u;
    for (uint32_t i = t; i < kWaves * radix; i += kThreads) s_cnt[(i / radix) * R + i % radix] = 0;
    {
        const uint32_t c0 = d0 < radix ? counts[(size_t)d0 * ntiles + tile] : 0u;
        const uint32_t c1 = d1 < radix ? counts[(size_t)d1 * ntiles + tile] : 0u;
        const uint32_t t0 = d0 < radix ? totals[d0] : 0u, t1 = d1 < radix ? totals[d1] : 0u;
        uint32_t all;
        const uint32_t digit_base = block_exclusive_scan(t0 + t1, s_wave, &all);
        if (d0 < radix) s_base[d0] = digit_base + c0;
        if (d1 < radix) s_base[d1] = digit_base + t0 + c1;
    }
    __syncthreads();
    rank_tile<ITEMS, BITS, R, kMaxDigitBits>(key, n, tile_base, first, shift, bits, s_cnt, rank);
    __syncthreads();
    const uint32_t tile_cnt0 = wave_offsets<R, kWaves>(d0, radix, s_cnt), tile_cnt1 = wave_offsets<R, kWaves>(d1, radix, s_cnt);
    {
        uint32_t all;
        const uint32_t st = block_exclusive_scan(tile_cnt0 + tile_cnt1, s_wave, &all);
        if (d0 < radix) s_start[d0] = st;
        if (d1 < radix) s_start[d1] = st + tile_cnt0;
    }
    __syncthreads();
    write_runs<ITEMS, R>(key, val, rank, n, tile_base, shift, mask, s_cnt, s_start, s_base, s_keys, s_vals, kout,
                         vout);
}

template <int ITEMS>
bool run_pass(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint32_t n, int shift,
              int bits, uint32_t* counts, uint32_t* totals, bool identity, hipStream_t s, bool counted)
{
    constexpr uint32_t TILE = kThreads * ITEMS;
    const uint32_t ntiles = (uint32_t)(((size_t)n + TILE - 1) / TILE);
    const uint32_t radix = 1u << bits;
    if (!counted)  // else the caller filled the count table while producing the keys
        hipLaunchKernelGGL(k_tile_counts<ITEMS>, dim3(ntiles), dim3(kThreads), 0, s, kin, n, shift, bits, counts,
                           ntiles);
    hipLaunchKernelGGL(k_scan_digits, dim3(radix), dim3(kThreads), 0, s, counts, ntiles, totals);
#define MFL_SCATTER(ID, B)                                                                                      \
    hipLaunchKernelGGL((k_scatter<ITEMS, ID, B>), dim3(ntiles), dim3(kThreads), 0, s, kin, vin, kout, vout, n, shift, \
                       bits, counts, totals, ntiles)
#define MFL_SCATTER_BITS(ID)                                                                                    \
    switch (bits) {                                                                                             \
    case 9: MFL_SCATTER(ID, 9); break;                                                                          \
    case 8: MFL_SCATTER(ID, 8); break;                                                                          \
    case 7: MFL_SCATTER(ID, 7); break;                                                                          \
    default: MFL_SCATTER(ID, 0); break;                                                                         \
    }
    if (identity) {
        MFL_SCATTER_BITS(true)
    } else {
        MFL_SCATTER_BITS(false)
    }
#undef MFL_SCATTER_BITS
#undef MFL_SCATTER
    return hip_ok(hipGetLastError(), "radix sort pass launch");
}

// Temporaries of one sort: key and value ping-pong buffers of n words each, the digit-major count
// table (radix x tiles) and the per-digit totals.
struct SortBuffers {
    uint32_t* tk[2];
    uint32_t* tv[2];
    uint32_t* counts;
    uint32_t* totals;
};

int pick_items(size_t n, int requested)
{
    if (requested == 4 || requested == 8 || requested == 16) return requested;
    if (n >= (size_t)8 << 20) return 16;
    if (n >= (size_t)1 << 20) return 8;
    return 4;
}

}  // namespace

RadixSort::~RadixSort()
{
    for (auto& b : tmp_keys_) b.release();
    for (auto& b : tmp_vals_) b.release();
    counts_.release();
    totals_.release();
}

int RadixSort::compile()
{
    int dev = 0;
    MFL_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MFL_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0) {
        set_error(std::string("libmfl kernels are built for gfx950, device is ") + prop.gcnArchName);
        return 0;
    }
    return 1;
}

int RadixSort::compute(uint32_t* keys, uint32_t* vals, size_t n, int start_bit, int end_bit, hipStream_t stream,
                       bool identity_vals)
{
    // argument checks of libCL/sort/oclRadixSort.cpp:64-80 (the 4,194,304 cap is lifted)
    if ((end_bit - start_bit) % 4 != 0) {
        set_error("end bit(" + std::to_string(end_bit) + ") - start bit(" + std::to_string(start_bit) +
                  ") must be divisible by 4");
        return 0;
    }
    return sort_bits(keys, vals, n, start_bit, end_bit, stream, identity_vals);
}

int RadixSort::reserve(size_t n, int nbits)
{
    if (n == 0 || nbits <= 0) return 1;
    const int passes = (nbits + kMaxDigitBits - 1) / kMaxDigitBits;
    const int items = pick_items(n, tuning.items);
    const size_t ntiles = (n + kThreads * items - 1) / (kThreads * items);
    const size_t bytes = n * sizeof(uint32_t);
    const int nbuf = passes == 1 ? 1 : 2;
    for (int i = 0; i < nbuf; i++)
        if (!tmp_keys_[i].fit(bytes) || !tmp_vals_[i].fit(bytes)) return 0;
    if (!counts_.fit(ntiles * kMaxRadix * sizeof(uint32_t)) || !totals_.fit(kMaxRadix * sizeof(uint32_t))) return 0;
    return 1;
}

void RadixSort::sig_append(std::string& sig) const
{
    for (const DeviceBuffer* b : {&tmp_keys_[0], &tmp_keys_[1], &tmp_vals_[0], &tmp_vals_[1], &counts_, &totals_}) {
        sig_add(sig, b->ptr);
        sig_add(sig, b->bytes);
    }
}

int RadixSort::first_pass_plan(size_t n, int start_bit, int end_bit, SortCountPlan* plan)
{
    const int nbits = end_bit - start_bit;
    if (n == 0 || n > 0xFFFFFFFFull || start_bit < 0 || end_bit > 32 || nbits <= 0) {
        set_error("radix sort: no count plan for this sort");
        return 0;
    }
    if (!reserve(n, nbits)) return 0;
    const int passes = (nbits + kMaxDigitBits - 1) / kMaxDigitBits;
    plan->items = pick_items(n, tuning.items);
    plan->tile_keys = (uint32_t)(kThreads * plan->items);
    plan->ntiles = (uint32_t)((n + plan->tile_keys - 1) / plan->tile_keys);
    plan->bits = (uint32_t)(nbits / passes + (nbits % passes ? 1 : 0));
    plan->counts = counts_.as<uint32_t>();
    return 1;
}

namespace {

// The passes of a sort of n pairs (1 <= n < 2^32, 0 <= start_bit < end_bit <= 32) over buffers
// that fit it.
int sort_passes(const SortBuffers& b, uint32_t* keys, uint32_t* vals, size_t n, int start_bit, int end_bit,
                hipStream_t stream, bool identity_vals, bool first_counted, int items)
{
    const int nbits = end_bit - start_bit;
    const int passes = (nbits + kMaxDigitBits - 1) / kMaxDigitBits;
    const size_t bytes = n * sizeof(uint32_t);
    const uint32_t* kin = keys;
    const uint32_t* vin = vals;
    int shift = start_bit;
    for (int p = 0; p < passes; p++) {
        const int bits = nbits / passes + (p < nbits % passes ? 1 : 0);
        const bool last = p == passes - 1 && passes > 1;
        uint32_t* kout = last ? keys : b.tk[p & 1];
        uint32_t* vout = last ? vals : b.tv[p & 1];
        const bool ident = identity_vals && p == 0;
        const bool counted = first_counted && p == 0;
        bool ok;
        switch (items) {
        case 16: ok = run_pass<16>(kin, vin, kout, vout, (uint32_t)n, shift, bits, b.counts, b.totals, ident, stream,
                                   counted); break;
        case 8: ok = run_pass<8>(kin, vin, kout, vout, (uint32_t)n, shift, bits, b.counts, b.totals, ident, stream,
                                 counted); break;
        default: ok = run_pass<4>(kin, vin, kout, vout, (uint32_t)n, shift, bits, b.counts, b.totals, ident, stream,
                                  counted); break;
        }
        if (!ok) return 0;
        kin = kout;
        vin = vout;
        shift += bits;
    }
    if (passes == 1) {
        MFL_HIP(hipMemcpyAsync(keys, b.tk[0], bytes, hipMemcpyDeviceToDevice, stream));
        MFL_HIP(hipMemcpyAsync(vals, b.tv[0], bytes, hipMemcpyDeviceToDevice, stream));
    }
    return 1;
}

// Checks shared by the sort entry points; 1 = go on with the sort, 0 = failed, -1 = nothing to sort.
int check_sort_args(const uint32_t* keys, const uint32_t* vals, size_t n, int start_bit, int end_bit)
{
    if (start_bit < 0 || end_bit > 32) {
        set_error("bit range must lie in [0, 32]");
        return 0;
    }
    if (n == 0 || start_bit >= end_bit) return -1;
    if (n > 0xFFFFFFFFull) {
        set_error("maximum sortable array size = 4294967295");
        return 0;
    }
    if (!keys || !vals) {
        set_error("radix sort: null key or value buffer");
        return 0;
    }
    return 1;
}

// ---- the sort without device allocation ------------------------------------------------------------
//
// compute_noalloc's temporaries live in `scratch`, a file-scope device array: each GPU holds its own copy
// from the library's load until the process ends. Layout (32-bit words): key ping-pong buffers at 0 and
// kNoAllocMaxPairs, value buffers at 2 and 3 x kNoAllocMaxPairs, the count table (at most
// kNoAllocMaxPairs / 1024 tiles of >= 1024 keys x 512 digits) at 4 x kNoAllocMaxPairs, the 512 digit
// totals after it. Every pass writes the count table and totals before reading them and every buffer
// before reading it, so nothing has to start cleared. Host code never uses the array's name as an
// address: it is a host-side stand-in; the device address comes from hipGetSymbolAddress.
constexpr size_t kNoAllocMaxPairs = MFL_SORT_NOALLOC_MAX_PAIRS;
constexpr size_t kNoAllocCounts = kNoAllocMaxPairs / (kThreads * 4) * kMaxRadix;
constexpr size_t kNoAllocWords = 4 * kNoAllocMaxPairs + kNoAllocCounts + kMaxRadix;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device ordinal, once looked up

// `scratch`'s address on the current device, looked up on the first call made with that device current.
uint32_t* scratch_address()
{
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device ordinal " + std::to_string(dev) + " above the no-allocation sort's device table");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return nullptr;
        g_scratch[dev] = static_cast<uint32_t*>(p);
    }
    return g_scratch[dev];
}

}  // namespace

int RadixSort::compute_noalloc(uint32_t* keys, uint32_t* vals, size_t n, int start_bit, int end_bit,
                               hipStream_t stream)
{
    if ((end_bit - start_bit) % 4 != 0) {  // libCL/sort/oclRadixSort.cpp:76-80, as compute()
        set_error("end bit(" + std::to_string(end_bit) + ") - start bit(" + std::to_string(start_bit) +
                  ") must be divisible by 4");
        return 0;
    }
    const int go = check_sort_args(keys, vals, n, start_bit, end_bit);
    if (go <= 0) return go < 0 ? 1 : 0;
    if (n > kNoAllocMaxPairs) {
        set_error("the no-allocation sort takes at most " + std::to_string(kNoAllocMaxPairs) + " pairs, not " +
                  std::to_string(n));
        return 0;
    }
    uint32_t* base = scratch_address();
    if (!base) return 0;
    const int items = pick_items(n, tuning.items);  // >= 4 keys per thread: ntiles <= kNoAllocMaxPairs / 1024
    SortBuffers b;
    b.tk[0] = base;
    b.tk[1] = base + kNoAllocMaxPairs;
    b.tv[0] = base + 2 * kNoAllocMaxPairs;
    b.tv[1] = base + 3 * kNoAllocMaxPairs;
    b.counts = base + 4 * kNoAllocMaxPairs;
    b.totals = b.counts + kNoAllocCounts;
    return sort_passes(b, keys, vals, n, start_bit, end_bit, stream, false, false, items);
}

int RadixSort::sort_bits(uint32_t* keys, uint32_t* vals, size_t n, int start_bit, int end_bit, hipStream_t stream,
                         bool identity_vals, bool first_counted)
{
    const int go = check_sort_args(keys, vals, n, start_bit, end_bit);
    if (go <= 0) return go < 0 ? 1 : 0;
    if (!reserve(n, end_bit - start_bit)) return 0;
    SortBuffers b;
    b.tk[0] = tmp_keys_[0].as<uint32_t>();
    b.tk[1] = tmp_keys_[1].as<uint32_t>();
    b.tv[0] = tmp_vals_[0].as<uint32_t>();
    b.tv[1] = tmp_vals_[1].as<uint32_t>();
    b.counts = counts_.as<uint32_t>();
    b.totals = totals_.as<uint32_t>();
    return sort_passes(b, keys, vals, n, start_bit, end_bit, stream, identity_vals, first_counted,
                       pick_items(n, tuning.items));
}

}  // namespace mfl
