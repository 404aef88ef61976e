// The 3D SPH step of oclFluid3D (libCL/phys/oclFluid3D.{h,cpp,cl}) as hand-written gfx950 kernels.
//
// Data layout in HBM is the reference's (libCL/phys/oclFluid3D.cpp:52-80): float4 AoS position,
// velocity, force in the caller's particle order; float4 sorted copies whose w lanes carry density
// (position) and pressure (velocity); uint32 cell keys and sorted indices; two uint32[2^18] bucket
// tables (cell start / end) that keep their contents between steps.
//
// Step (oclFluid3D::compute, cpp:293-342):
//   k_hash_tiles    cell key per particle + reset of the first clear_count buckets (cl:57-64,
//                   cpp:309-313; nothing between the two reads the tables, so the reset moves to the
//                   start of the step), each tile of keys then ordered by its high digit in LDS
//   k_bucket_runs   the rest of the stable key/index sort over bits [0, 18) (cpp:299; keys < 2^18):
//                   one block per high digit assembles that digit's run from the tiles, orders it by
//                   bits [0, 9) and writes the inverse permutation and the start / end of each occupied
//                   bucket (cl:85-113) from the same histograms; up to 2^14 particles the hash kernel
//                   (k_hash_keys) leaves the keys in caller order and each block selects its own
//   k_scatter_sorted sorted pos/vel copy, written from the caller's order   (cl:66-72)
//   density pass    27-bucket neighbourhood, poly6, pressure                (cl:150-200)
//   force pass      27-bucket pressure + viscosity forces, clip box, clamp, Euler, gravity
//                   (cl:202-255, 315-361, 258-271, 294-298, 284-292 fused; each particle's update
//                    only touches its own force/velocity/position, so fusion is exact)
// With an EVT_INTEGRATE callback the force pass runs alone and the callback follows. The reference's
// kernels one by one (getKernel + enqueue) are k_hash, k_reorder_clear, k_find_bounds, k_density,
// k_forces and the integration kernels below.
#include "fluid3d.hpp"
#include "sph_device.hpp"
#include "tile.hpp"
#include "tile_rank.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace mfl {
namespace {

using namespace sph;

template <bool WRITE_INDEX, bool CLEAR>
__global__ __launch_bounds__(kThreads) void k_hash(const float4* __restrict__ pos, uint32_t n, float h,
                                                   uint32_t* __restrict__ cell, uint32_t* __restrict__ index,
                                                   mfl_params* __restrict__ params_buf, mfl_params p,
                                                   uint32_t* __restrict__ cell_start, uint32_t* __restrict__ cell_end,
                                                   uint32_t clear)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i == 0 && params_buf) *params_buf = p;  // keeps getParamBuffer() in step with the host Params
    if (CLEAR && i < clear) {
        cell_start[i] = kEmpty;
        cell_end[i] = kEmpty;
    }
    if (i >= n) return;
    const float4 q = pos[i];
    cell[i] = cell_hash(grid_coord(q.x, h), grid_coord(q.y, h), grid_coord(q.z, h));
    if (WRITE_INDEX) index[i] = i;
}

// The step's hash kernel for at most kSelectMax particles: cell key per particle (cl:57-64) in caller
// order, the reset of the first `clear` buckets (cpp:309-313; nothing between the two reads the tables,
// so the reset moves to the start of the step) and the parameter copy. ITEMS keys per thread, every
// load issued before the first hash.
template <int ITEMS>
__global__ __launch_bounds__(kThreads) void k_hash_keys(const float4* __restrict__ pos, uint32_t n, float h,
                                                        uint32_t* __restrict__ cell, mfl_params* __restrict__ params_buf,
                                                        mfl_params p, uint32_t* __restrict__ cell_start,
                                                        uint32_t* __restrict__ cell_end, uint32_t clear)
{
    if (blockIdx.x == 0 && threadIdx.x == 0 && params_buf) *params_buf = p;  // getParamBuffer() in step
    for (uint32_t c = blockIdx.x * kThreads + threadIdx.x; c < clear; c += gridDim.x * kThreads) {
        cell_start[c] = kEmpty;
        cell_end[c] = kEmpty;
    }
    const size_t base = (size_t)blockIdx.x * (kThreads * ITEMS);
    float4 q[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const size_t i = base + (size_t)k * kThreads + threadIdx.x;
        q[k] = pos[i < n ? i : n - 1];  // a clamped index keeps the loads unconditional
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const size_t i = base + (size_t)k * kThreads + threadIdx.x;
        if (i < n) cell[i] = cell_hash(grid_coord(q[k].x, h), grid_coord(q[k].y, h), grid_coord(q[k].z, h));
    }
}

// The step's hash kernel for more than kSelectMax particles: k_hash_keys's keys, table reset and
// parameter copy, and then the tile (kThreads x ITEMS particles, sort tile xcd_tile()) ordered stably by
// the keys' high digit (bits [9, 18)) in LDS and written back in place: tkeys / tvals[tile * TILE + p]
// hold the tile's keys and their particle indices in (high digit, index) order, counts[d ntiles + tile]
// the tile's number of keys of digit d and starts[d ntiles + tile] where they begin in the tile. Item r
// of wave w, lane l is particle tile * TILE + (w ITEMS + r) 64 + l ((wave, round, lane) order is input
// order, as tile_rank.hpp ranks).
template <int ITEMS>
__global__ __launch_bounds__(kThreads) void k_hash_tiles(const float4* __restrict__ pos, uint32_t n, float h,
                                                         mfl_params* __restrict__ params_buf, mfl_params p,
                                                         uint32_t* __restrict__ cell_start,
                                                         uint32_t* __restrict__ cell_end, uint32_t clear,
                                                         uint32_t* __restrict__ tkeys, uint32_t* __restrict__ tvals,
                                                         uint32_t* __restrict__ counts, uint32_t* __restrict__ starts,
                                                         uint32_t ntiles)
{
    constexpr int TILE = kThreads * ITEMS;
    constexpr uint32_t R = 512;
    constexpr int WAVES = kThreads / kWave;
    __shared__ uint32_t s_cnt[WAVES * R];
    __shared__ uint32_t s_start[R];
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_wave[WAVES];
    const uint32_t t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
    for (uint32_t i = t; i < WAVES * R; i += kThreads) s_cnt[i] = 0;
    if (blockIdx.x == 0 && t == 0 && params_buf) *params_buf = p;  // getParamBuffer() in step
    for (uint32_t c = blockIdx.x * kThreads + t; c < clear; c += gridDim.x * kThreads) {
        cell_start[c] = kEmpty;
        cell_end[c] = kEmpty;
    }
    const uint32_t tile = xcd_tile();
    const size_t tile_base = (size_t)tile * TILE;
    const uint32_t first = (wave * ITEMS) * kWave + lane;
    // every load issued before the first hash (a clamped index keeps the loads unconditional)
    float4 q[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const size_t i = tile_base + first + (size_t)r * kWave;
        q[r] = pos[i < n ? i : n - 1];
    }
    uint32_t key[ITEMS], rank[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) key[r] = cell_hash(grid_coord(q[r].x, h), grid_coord(q[r].y, h), grid_coord(q[r].z, h));
    __syncthreads();
    rank_tile<ITEMS, 9, R, 9>(key, n, tile_base, first, 9, 9, s_cnt, rank);
    __syncthreads();
    const uint32_t d0 = 2u * t, d1 = d0 + 1u;  // thread t owns high digits 2t and 2t + 1
    const uint32_t c0 = wave_offsets<R, WAVES>(d0, R, s_cnt), c1 = wave_offsets<R, WAVES>(d1, R, s_cnt);
    {
        uint32_t all;
        const uint32_t st = block_scan<kThreads>(c0 + c1, s_wave, &all);
        s_start[d0] = st;
        s_start[d1] = st + c0;
        counts[(size_t)d0 * ntiles + tile] = c0;
        counts[(size_t)d1 * ntiles + tile] = c1;
        starts[(size_t)d0 * ntiles + tile] = st;
        starts[(size_t)d1 * ntiles + tile] = st + c0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const size_t i = tile_base + first + (size_t)r * kWave;
        if (i < n) {
            const uint32_t d = key[r] >> 9;
            const uint32_t at = s_start[d] + s_cnt[wave * R + d] + rank[r];
            s_keys[at] = key[r];
            s_vals[at] = (uint32_t)i;
        }
    }
    __syncthreads();
    const size_t remain = n - tile_base;
    const uint32_t tile_n = remain < (size_t)TILE ? (uint32_t)remain : (uint32_t)TILE;
    for (uint32_t at = t; at < tile_n; at += kThreads) {
        tkeys[tile_base + at] = s_keys[at];
        tvals[tile_base + at] = s_vals[at];
    }
}

// Sorted gather (cl:66-72) over [0, n) and the per-step bucket reset (cpp:309-313) over [0, clear).
__global__ __launch_bounds__(kThreads) void k_reorder_clear(const uint32_t* __restrict__ index,
                                                            const float4* __restrict__ pos,
                                                            const float4* __restrict__ vel,
                                                            float4* __restrict__ spos, float4* __restrict__ svel,
                                                            uint32_t n, uint32_t* __restrict__ cell_start,
                                                            uint32_t* __restrict__ cell_end, uint32_t clear,
                                                            uint32_t* __restrict__ inv)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        const uint32_t s = index[i];
        if (s < n) {  // index is a permutation of [0, n)
            spos[i] = pos[s];
            svel[i] = vel[s];
            if (inv) inv[s] = i;
        }
    }
    if (i < clear) {
        cell_start[i] = kEmpty;
        cell_end[i] = kEmpty;
    }
}

// Particle count from which the reorder stage copies the state in the caller's order (two launches):
// 2^18 particles hold 8 MiB of pos + vel, twice an XCD's L2.
constexpr uint32_t kScatterCopyMin = 1u << 18;

// Largest particle count whose run blocks select their keys from the hash kernel's caller-order keys:
// each of the 512 blocks reads all of them (64 KiB, from L2) and takes its high digit's own, which at
// 16,384 particles costs the step a little less than ordering the hash kernel's tiles by the high
// digit and assembling each run from the tiles' segments (37.1 against 37.5 µs per step).
constexpr uint32_t kSelectMax = 1u << 14;

// Second half of the step's sort, one block per high digit c (bits [9, 18) of the keys, which are
// < 2^18; blocks dealt to the XCDs in contiguous ranges): the block assembles the run of keys with high
// digit c in input order and orders it stably by bits [0, 9): chunk by chunk in input order, ranks by
// wave ballot matching (tile_rank.hpp's scheme), the chunk reordered by digit in LDS and written out in
// runs at each digit's running offset (a run longer than one chunk takes a digit histogram first; one
// chunk long, as usual, its histogram is the chunk's). High digit c then low digit d is bucket
// c * 512 + d, so the run holds every key of buckets c * 512 .. c * 512 + 511, and the histogram gives
// their bounds directly (cl:85-113: start and end written for occupied buckets only). The block also
// writes each particle's sorted slot (inv) and, GATHER, copies the sorted state by gathering through the
// index (for steps whose pos / vel sit in L2 anyway; larger steps copy with k_scatter_sorted).
// MODE kRunsTiles: k_hash_tiles left each tile ordered by high digit; the run is the concatenation, in
// tile order, of every tile's digit-c segment (counts / starts give each segment's length and place),
// and its base is the sum over the tiles of their keys below digit c, which is the sum of the segments'
// starts. MODE kRunsSelect (at most kSelectMax particles): the keys in caller order, selected below.
constexpr int kRunThreads = 512;     // one thread per low digit
constexpr int kRunWaves = kRunThreads / kWave;
constexpr int kRunItems = 8;         // keys per thread and chunk
constexpr uint32_t kRunDigits = 512;
constexpr uint32_t kRunChunk = kRunThreads * kRunItems;
enum : int { kRunsTiles = 0, kRunsSelect = 1 };

template <bool GATHER, int MODE>
__global__ __launch_bounds__(kRunThreads) void k_bucket_runs(const uint32_t* __restrict__ kin,
                                                             const uint32_t* __restrict__ vin,
                                                             const uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ starts, uint32_t ntiles,
                                                             uint32_t tile_keys,
                                                             uint32_t* __restrict__ cell, uint32_t* __restrict__ index,
                                                             uint32_t* __restrict__ cell_start,
                                                             uint32_t* __restrict__ cell_end, uint32_t* __restrict__ inv,
                                                             const float4* __restrict__ pos,
                                                             const float4* __restrict__ vel, float4* __restrict__ spos,
                                                             float4* __restrict__ svel)
{
    constexpr bool SELECT = MODE == kRunsSelect;
    __shared__ uint32_t run_off[kRunDigits];          // histogram, then each digit's next slot in the run
    __shared__ uint32_t chunk_start[kRunDigits];      // each digit's first slot in the chunk's order
    __shared__ uint32_t s_cnt[kRunWaves * kRunDigits];  // a chunk's per-wave digit counts
    __shared__ uint32_t s_keys[kRunChunk];            // the chunk ordered by digit
    __shared__ uint32_t s_vals[kRunChunk];
    __shared__ uint32_t s_wave[kRunWaves];
    const uint32_t c = xcd_tile();
    const uint32_t t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;  // thread t owns low digit t
    const uint64_t lower_lanes = (1ull << lane) - 1ull;
    // MODE kRunsTiles: thread t's tiles are tt0 .. tt1 - 1, its first segment starts at seg_off in the
    // run; tiles_each(lo, hi, f) calls f(run position, key index in tkeys / tvals) for the run positions
    // in [lo, hi) that its segments hold, four loads in flight at a time
    const uint32_t tpt = MODE == kRunsTiles ? (ntiles + kRunThreads - 1) / kRunThreads : 0u;
    const uint32_t tt0 = min(t * tpt, ntiles), tt1 = min(tt0 + tpt, ntiles);
    uint32_t seg_off = 0;
    uint32_t ct0 = 0, st0 = 0;  // the first tile's count and start, kept from the setup
    uint32_t pk[4] = {0u, 0u, 0u, 0u}, pv[4] = {0u, 0u, 0u, 0u};  // its first four entries, loaded there too
    auto tiles_each = [&](uint32_t lo, uint32_t hi, auto&& f) {
        uint32_t off = seg_off;
        for (uint32_t tt = tt0; tt < tt1; tt++) {
            const uint32_t ct = tt == tt0 ? ct0 : counts[(size_t)c * ntiles + tt];
            const size_t at = (size_t)tt * tile_keys + (tt == tt0 ? st0 : starts[(size_t)c * ntiles + tt]);
            const uint32_t i_lo = lo > off ? min(lo - off, ct) : 0u, i_hi = hi > off ? min(hi - off, ct) : 0u;
            for (uint32_t i = i_lo; i < i_hi; i += 4) {
                uint32_t kk[4], vv[4];
                const bool pre = tt == tt0 && i == 0u;  // entries 0-3 of the first tile: prefetched
#pragma unroll
                for (uint32_t r = 0; r < 4; r++) {
                    const size_t j = at + min(i + r, i_hi - 1u);
                    kk[r] = pre ? pk[r] : kin[j];
                    vv[r] = pre ? pv[r] : vin[j];
                }
#pragma unroll
                for (uint32_t r = 0; r < 4; r++)
                    if (i + r < i_hi) f(off + i + r, kk[r], vv[r]);
            }
            off += ct;
        }
    };
    // SELECT: the keys are the hash kernel's, in input order (kin[0 .. n), n <= kSelectMax, followed by
    // 0xFFFFFFFF up to kSelectMax, which no digit matches); every block reads all of them and takes
    // those of its high digit. Wave w reads the contiguous keys w kSelectPer .. (w + 1) kSelectPer - 1
    // (round r, lane l: key w kSelectPer + 64 r + l) in batches of 8 rounds, the next batch's loads in
    // flight while one is counted, so a match's rank in input order is the matches of the earlier waves
    // plus those of the earlier rounds and lower lanes of its own wave. The count also stages each wave's
    // first kSelectStage matches, in order, in its own slice of s_keys / s_vals, and builds the run's
    // low-digit histogram in run_off; after one exchange of the wave totals the staged matches move to
    // their ranks. A wave with more matches, or a run longer than a chunk, takes sel_store instead: a
    // second read of the keys that stores the matches of ranks skip .. skip + kRunChunk - 1.
    constexpr uint32_t kSelectPer = kSelectMax / kRunWaves, kSelectRounds = kSelectPer / kWave;
    constexpr uint32_t kSelectBatch = 8, kSelectStage = kRunChunk / kRunWaves;
    __shared__ uint32_t s_sel[2 * kRunWaves];
    const uint32_t* src = kin + wave * kSelectPer + lane;
    auto sel_store = [&](uint32_t first_rank, uint32_t skip) {
        uint32_t rank = first_rank;
        for (uint32_t b = 0; b < kSelectRounds; b += kSelectBatch) {
            uint32_t k[kSelectBatch];
#pragma unroll
            for (uint32_t r = 0; r < kSelectBatch; r++) k[r] = src[(b + r) * kWave];
#pragma unroll
            for (uint32_t r = 0; r < kSelectBatch; r++) {
                const bool m = (k[r] >> 9) == c;
                const uint64_t bal = __ballot(m);
                const uint32_t at = rank + (uint32_t)__popcll(bal & lower_lanes);
                if (m && at >= skip && at - skip < kRunChunk) {
                    s_keys[at - skip] = k[r];
                    s_vals[at - skip] = wave * kSelectPer + (b + r) * kWave + lane;
                }
                rank += (uint32_t)__popcll(bal);
            }
        }
    };
    uint32_t base = 0, cnt = 0, first_rank = 0;
    bool staged = true;  // every match of the run in s_keys / s_vals after the select (else sel_store)
    if constexpr (SELECT) {
        run_off[t] = 0;
        __syncthreads();
        uint32_t mine = 0, below = 0;  // the wave's matches so far (uniform); keys below digit c (per lane)
        uint32_t ka[kSelectBatch], kb[kSelectBatch];
#pragma unroll
        for (uint32_t r = 0; r < kSelectBatch; r++) ka[r] = src[r * kWave];
        for (uint32_t b = 0; b < kSelectRounds; b += kSelectBatch) {
            if (b + kSelectBatch < kSelectRounds) {
#pragma unroll
                for (uint32_t r = 0; r < kSelectBatch; r++) kb[r] = src[(b + kSelectBatch + r) * kWave];
            }
#pragma unroll
            for (uint32_t r = 0; r < kSelectBatch; r++) {
                const uint32_t hi = ka[r] >> 9;
                const bool m = hi == c;
                below += hi < c ? 1u : 0u;
                const uint64_t bal = __ballot(m);
                if (m) {
                    const uint32_t at = mine + (uint32_t)__popcll(bal & lower_lanes);
                    if (at < kSelectStage) {
                        s_keys[wave * kSelectStage + at] = ka[r];
                        s_vals[wave * kSelectStage + at] = wave * kSelectPer + (b + r) * kWave + lane;
                    }
                    atomicAdd(&run_off[ka[r] & (kRunDigits - 1u)], 1u);
                }
                mine += (uint32_t)__popcll(bal);
            }
#pragma unroll
            for (uint32_t r = 0; r < kSelectBatch; r++) ka[r] = kb[r];
        }
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) below += (uint32_t)__shfl_xor((int)below, o, kWave);
        if (lane == 0) {
            s_sel[wave] = mine;
            s_sel[kRunWaves + wave] = below;
        }
        __syncthreads();
        base = cnt = 0;
        uint32_t most = 0;
#pragma unroll
        for (uint32_t w = 0; w < kRunWaves; w++) {
            const uint32_t m = s_sel[w];
            first_rank += w < wave ? m : 0u;
            cnt += m;
            base += s_sel[kRunWaves + w];
            most = max(most, m);
        }
        if (cnt == 0) return;  // block-uniform
        staged = most <= kSelectStage && cnt <= kRunChunk;
        if (staged) {  // the staged matches to their ranks (every slice read before any is written)
            uint32_t mk[kSelectStage / kWave], mv[kSelectStage / kWave];
#pragma unroll
            for (uint32_t i = 0; i < kSelectStage / kWave; i++) {
                const uint32_t j = i * kWave + lane;
                mk[i] = j < mine ? s_keys[wave * kSelectStage + j] : 0u;
                mv[i] = j < mine ? s_vals[wave * kSelectStage + j] : 0u;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t i = 0; i < kSelectStage / kWave; i++) {
                const uint32_t j = i * kWave + lane;
                if (j < mine) {
                    s_keys[first_rank + j] = mk[i];
                    s_vals[first_rank + j] = mv[i];
                }
            }
        }
        __syncthreads();  // run_off complete, s_sel read, the staged chunk in place
    } else {
        // this thread's tiles: a contiguous range, so that its segments follow each other in the run
        uint32_t mine = 0, below = 0;
        for (uint32_t tt = tt0; tt < tt1; tt++) {
            const uint32_t ct = counts[(size_t)c * ntiles + tt], st = starts[(size_t)c * ntiles + tt];
            if (tt == tt0) {
                ct0 = ct;
                st0 = st;
                if (ct > 0u) {  // the segment's first entries, in flight during the block scans below
                    const size_t at = (size_t)tt * tile_keys + st;
#pragma unroll
                    for (uint32_t r = 0; r < 4; r++) {
                        pk[r] = kin[at + min(r, ct - 1u)];
                        pv[r] = vin[at + min(r, ct - 1u)];
                    }
                }
            }
            mine += ct;
            below += st;
        }
        uint32_t all;
        seg_off = block_scan<kRunThreads>(mine, s_wave, &cnt);
        (void)block_scan<kRunThreads>(below, s_wave, &all);
        base = all;
        run_off[t] = 0;
        __syncthreads();
        if (cnt == 0) return;  // block-uniform
    }
    // bounds of the run's buckets from its digit histogram, whose exclusive scan starts each digit's slots
    // (block-uniform call); returns the scan
    auto bounds = [&](uint32_t h) {
        uint32_t all;
        const uint32_t st = block_scan<kRunThreads>(h, s_wave, &all);
        if (h) {
            cell_start[(c << 9) | t] = base + st;
            cell_end[(c << 9) | t] = base + st + h;
        }
        run_off[t] = st;
        return st;
    };
    // a run of one chunk takes its histogram from the chunk's ranking below; a longer one counts first
    // (SELECT counted it while selecting)
    const bool one = !SELECT && cnt <= kRunChunk;
    if (SELECT) {
        bounds(run_off[t]);
    } else if (!one) {
        tiles_each(0u, cnt, [&](uint32_t, uint32_t k, uint32_t) { atomicAdd(&run_off[k & (kRunDigits - 1u)], 1u); });
        __syncthreads();
        bounds(run_off[t]);
    }
    for (uint32_t c0 = 0; c0 < cnt; c0 += kRunChunk) {
        for (uint32_t i = t; i < kRunWaves * kRunDigits; i += kRunThreads) s_cnt[i] = 0;
        if constexpr (SELECT) {  // this chunk's matches into s_keys / s_vals, unless staged there
            if (!staged) sel_store(first_rank, c0);
        } else {  // this chunk's part of every tile segment
            tiles_each(c0, c0 + kRunChunk, [&](uint32_t q, uint32_t k, uint32_t v) {
                s_keys[q - c0] = k;
                s_vals[q - c0] = v;
            });
        }
        __syncthreads();
        // item r of wave w, lane l is key c0 + (w kRunItems + r) 64 + l: (wave, round, lane) order is input order
        uint32_t key[kRunItems], val[kRunItems], rank[kRunItems];
#pragma unroll
        for (int r = 0; r < kRunItems; r++) {
            const uint32_t q = c0 + (wave * kRunItems + r) * kWave + lane;
            const bool valid = q < cnt;
            key[r] = valid ? s_keys[q - c0] : 0u;  // the chunk, gathered into LDS above
            val[r] = valid ? s_vals[q - c0] : 0u;
            rank[r] = 0;
        }
        __syncthreads();  // read before the reorder below overwrites s_keys / s_vals
#pragma unroll
        for (int r = 0; r < kRunItems; r++) {
            // rounds wholly past the run's end rank nothing (wave-uniform: the rest of the wave's rounds too)
            if (c0 + (uint32_t)(wave * kRunItems + r) * kWave >= cnt) break;
            const uint32_t q = c0 + (wave * kRunItems + r) * kWave + lane;
            const bool valid = q < cnt;
            const uint32_t d = key[r] & (kRunDigits - 1u);
            // lanes whose digit differs in some bit (or that hold no key), per bit b: the ballot of bit b
            // XOR this lane's bit b replicated
            const uint64_t live = __ballot(valid);
            uint32_t differ_lo = ~(uint32_t)live, differ_hi = ~(uint32_t)(live >> 32);
#pragma unroll
            for (int b = 0; b < 9; b++) {
                const uint32_t rep = (uint32_t)((int)(d << (31 - b)) >> 31);
                const uint64_t bal = __ballot(rep != 0u);
                differ_lo |= (uint32_t)bal ^ rep;
                differ_hi |= (uint32_t)(bal >> 32) ^ rep;
            }
            const uint64_t same = ~((uint64_t)differ_hi << 32 | differ_lo);
            const int leader = __ffsll((unsigned long long)same) - 1;
            uint32_t before = 0;
            if (valid && (int)lane == leader) {
                before = s_cnt[wave * kRunDigits + d];
                s_cnt[wave * kRunDigits + d] = before + (uint32_t)__popcll(same);
            }
            before = __shfl(before, leader & (kWave - 1), kWave);
            rank[r] = before + (uint32_t)__popcll(same & lower_lanes);
        }
        __syncthreads();
        uint32_t mine = 0;  // the chunk's keys of digit t
#pragma unroll
        for (int w = 0; w < kRunWaves; w++) {
            const uint32_t a = s_cnt[w * kRunDigits + t];
            s_cnt[w * kRunDigits + t] = mine;
            mine += a;
        }
        if (one) {
            chunk_start[t] = bounds(mine);  // the chunk is the whole run
        } else {
            uint32_t all;
            chunk_start[t] = block_scan<kRunThreads>(mine, s_wave, &all);
        }
        __syncthreads();
        // the chunk in digit order in LDS, then written out in runs
#pragma unroll
        for (int r = 0; r < kRunItems; r++) {
            const uint32_t q = c0 + (wave * kRunItems + r) * kWave + lane;
            if (q < cnt) {
                const uint32_t d = key[r] & (kRunDigits - 1u);
                const uint32_t p = chunk_start[d] + s_cnt[wave * kRunDigits + d] + rank[r];
                s_keys[p] = key[r];
                s_vals[p] = val[r];
            }
        }
        __syncthreads();
        const uint32_t chunk_n = min(cnt - c0, kRunChunk);
        for (uint32_t p = t; p < chunk_n; p += kRunThreads) {
            const uint32_t k = s_keys[p], v = s_vals[p];
            const uint32_t d = k & (kRunDigits - 1u);
            const uint32_t o = base + run_off[d] + (p - chunk_start[d]);
            cell[o] = k;
            index[o] = v;
            inv[v] = o;
            if (GATHER) {
                spos[o] = pos[v];
                svel[o] = vel[v];
            }
        }
        __syncthreads();
        run_off[t] += mine;
    }
}

__global__ __launch_bounds__(kThreads) void k_fill_buckets(uint32_t* __restrict__ cell_start,
                                                           uint32_t* __restrict__ cell_end, uint32_t count)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < count) {
        cell_start[i] = kEmpty;
        cell_end[i] = kEmpty;
    }
}

// cl:85-113. Every occupied bucket gets both its start and its end written; empty buckets keep
// whatever they held (0xFFFFFFFF if reset this step, a stale range otherwise).
__global__ __launch_bounds__(kThreads) void k_find_bounds(const uint32_t* __restrict__ cell, uint32_t n,
                                                          uint32_t* __restrict__ cell_start,
                                                          uint32_t* __restrict__ cell_end)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t h = cell[i] & (kBuckets - 1u);  // keys are < 2^18 by construction
    if (i == 0) {
        cell_start[h] = 0;
        return;
    }
    const uint32_t prev = cell[i - 1] & (kBuckets - 1u);
    if (h != prev) {
        cell_end[prev] = i;
        cell_start[h] = i;
    }
    if (i == n - 1) cell_end[h] = n;
}

__global__ __launch_bounds__(kThreads) void k_density(float4* __restrict__ spos, float4* __restrict__ svel,
                                                      const uint32_t* __restrict__ cell_start,
                                                      const uint32_t* __restrict__ cell_end, uint32_t n,
                                                      mfl_params p, uint32_t* __restrict__ counts_out)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float h = p.cellSize;
    const float h2_hi = h * h * 1.0000005f;
    const float4 pi = spos[i];
    const int gx = grid_coord(pi.x, h), gy = grid_coord(pi.y, h), gz = grid_coord(pi.z, h);
    float acc = 0.f;
    int count = 0;
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                const uint32_t c = cell_hash(gx + x, gy + y, gz + z);
                const uint32_t start = cell_start[c];
                if (start == kEmpty) continue;
                const uint32_t end = min(cell_end[c], n);  // stale or written ranges always lie in [0, n]
                for (uint32_t j = start; j < end; j++) {
                    if (j == i) continue;
                    const float4 pj = spos[j];
                    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    float vlen;
                    if (within(dist2(dx, dy, dz), h, h2_hi, &vlen)) {
                        const float t = h * h - vlen * vlen;
                        acc += t * t * t;
                        count++;
                    }
                }
            }
    if (counts_out) counts_out[i] = (uint32_t)count;
    float rho, prs;
    if (count > 2) {
        rho = fmaxf(1.0f, p.mass * p.kernelConstant * acc);
        prs = p.pressure + p.stiffness * (rho - p.density);
    } else {
        rho = p.density;
        prs = p.pressure;
    }
    spos[i].w = rho;
    svel[i].w = prs;
}

__global__ __launch_bounds__(kThreads) void k_forces(const float4* __restrict__ spos, const float4* __restrict__ svel,
                                                     float4* __restrict__ force, const uint32_t* __restrict__ index,
                                                     const uint32_t* __restrict__ cell_start,
                                                     const uint32_t* __restrict__ cell_end, uint32_t n, mfl_params p)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float4 pi = spos[i], vi = svel[i];
    float3 P, V;
    sph_force_sum(spos, svel, cell_start, cell_end, i, n, pi, vi, p.cellSize, P, V);
    const uint32_t o = index[i];
    if (o >= n) return;  // index is a permutation of [0, n)
    force[o] = add_sph_force(force[o], P, V, p);
}

// ---- neighbour-list path of compute() -----------------------------------------------------------
//
// The density pass walks the reference's 27-bucket neighbourhood exactly (z, y, x order, ascending j,
// duplicate buckets and stale ranges included) and records every accepted visit j of particle i in a
// slot-major list nbr[k * n + i], k < count[i] <= kNbrCap. The force pass uses the same acceptance
// (positions do not change between the two passes), so it sums over the recorded visits instead of
// re-testing ~27 buckets of candidates; a particle with more than kNbrCap visits falls back to the
// full traversal. Visits repeated by the reference (a bucket appearing twice among the 27) appear
// twice in the list, so multiplicities are kept.

// Traversal order of the neighbour passes. Results do not depend on it (each particle's sums are
// computed by one thread in a fixed visit order); it only decides which particles share a wave, an
// L1 and an XCD's L2. Hash order scatters neighbouring cells over the whole sorted array, so the
// passes walk particles in caller order or in a periodically refreshed Morton order of their cells
// and reach their sorted slot through the inverse permutation written by the reorder stage.
enum : int { kOrderSorted = 0, kOrderCaller = 1, kOrderListed = 2 };

template <int ORDER>
__device__ inline uint32_t sorted_slot(uint32_t t, uint32_t n, const uint32_t* __restrict__ porder,
                                       const uint32_t* __restrict__ inv)
{
    if (ORDER == kOrderSorted) return t;
    if (ORDER == kOrderCaller) return inv[t];
    const uint32_t o = porder[t];
    return o < n ? inv[o] : n;
}

template <int ORDER>
__device__ inline uint32_t caller_slot(uint32_t t, uint32_t i, const uint32_t* __restrict__ porder,
                                       const uint32_t* __restrict__ index)
{
    if (ORDER == kOrderSorted) return index[i];
    if (ORDER == kOrderCaller) return t;
    return porder[t];
}

template <int ORDER>
__global__ __launch_bounds__(kThreads) void k_density_nbr(float4* __restrict__ spos, float4* __restrict__ svel,
                                                          const uint32_t* __restrict__ cell_start,
                                                          const uint32_t* __restrict__ cell_end, uint32_t n,
                                                          mfl_params p, uint32_t* __restrict__ nbr,
                                                          uint32_t* __restrict__ ncount,
                                                          const uint32_t* __restrict__ porder,
                                                          const uint32_t* __restrict__ inv,
                                                          uint32_t* __restrict__ counts_out)
{
    const uint32_t t = xcd_block() * kThreads + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = sorted_slot<ORDER>(t, n, porder, inv);
    if (i >= n) return;
    const float h = p.cellSize;
    const float h2 = h * h;
    const float h2_lo = h2 * 0.9999996f, h2_hi = h2 * 1.0000005f;
    const float4 pi = spos[i];
    const int gx = grid_coord(pi.x, h), gy = grid_coord(pi.y, h), gz = grid_coord(pi.z, h);
    float acc = 0.f;
    uint32_t count = 0;
    for (int z = -1; z <= 1; z++) {
        // the nine bucket ranges of one z slab are fetched together (one memory latency, not nine)
        uint32_t st[9], en[9];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const uint32_t c = cell_hash(gx + (k % 3) - 1, gy + (k / 3) - 1, gz + z);
            st[k] = cell_start[c];
            en[k] = cell_end[c];
        }
#pragma unroll
        for (int k = 0; k < 9; k++) {
            if (st[k] == kEmpty) continue;
            const uint32_t end = min(en[k], n);
            for (uint32_t j = st[k]; j < end; j++) {
                const float4 pj = spos[j];
                const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                const float r2 = dist2(dx, dy, dz);
                if (j != i && accept_r2(r2, h, h2_lo, h2_hi)) {
                    const float w = h2 - r2;  // h^2 - vlen^2 of cl:123
                    acc = fmaf(w * w, w, acc);
                    if (count < kNbrCap) nbr[nbr_at(t, count, n)] = j;
                    count++;
                }
            }
        }
    }
    float rho, prs;
    if (count > 2) {
        rho = fmaxf(1.0f, p.mass * p.kernelConstant * acc);
        prs = p.pressure + p.stiffness * (rho - p.density);
    } else {
        rho = p.density;
        prs = p.pressure;
    }
    spos[i].w = rho;
    svel[i].w = prs;
    ncount[t] = count;
    if (counts_out) counts_out[i] = count;
}

template <int ORDER>
__global__ __launch_bounds__(kThreads) void k_density_grp(float4* __restrict__ spos, float4* __restrict__ svel,
                                                          const uint32_t* __restrict__ cell_start,
                                                          const uint32_t* __restrict__ cell_end, uint32_t n,
                                                          mfl_params p, uint32_t* __restrict__ nbr,
                                                          uint32_t* __restrict__ ncount,
                                                          const uint32_t* __restrict__ porder,
                                                          const uint32_t* __restrict__ inv,
                                                          uint32_t* __restrict__ counts_out)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t grp = lane / kGroupLanes, slot = lane % kGroupLanes;
    const uint32_t wave = (xcd_block() * kThreads + threadIdx.x) / kWave;
    const uint32_t t = wave * (kWave / kGroupLanes) + grp;
    uint32_t i = t < n ? sorted_slot<ORDER>(t, n, porder, inv) : n;
    const bool live = i < n;
    const float h = p.cellSize;
    const float h2 = h * h;
    const float h2_lo = h2 * 0.9999996f, h2_hi = h2 * 1.0000005f;
    const float4 pi = live ? spos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    int gx = grid_coord(pi.x, h), gy = grid_coord(pi.y, h), gz = grid_coord(pi.z, h);
    // the first live lane's cell; the wave is uniform when every live lane shares it
    const uint64_t live_mask = __ballot(live);
    if (live_mask == 0) return;
    const int src = __ffsll((unsigned long long)live_mask) - 1;
    const int gx0 = __shfl(gx, src, kWave), gy0 = __shfl(gy, src, kWave), gz0 = __shfl(gz, src, kWave);
    const bool same = !live || (gx == gx0 && gy == gy0 && gz == gz0);
    float acc = 0.f;
    uint32_t count = 0;
    const uint32_t grp_shift = grp * kGroupLanes;
    if (__ballot(!same) == 0) {
        density_visits(__builtin_amdgcn_readfirstlane(gx0), __builtin_amdgcn_readfirstlane(gy0),
                       __builtin_amdgcn_readfirstlane(gz0), pi, i, t, slot, grp_shift, live, spos, cell_start,
                       cell_end, n, h, h2, h2_lo, h2_hi, nbr, n, acc, count);
    } else {
        density_visits(gx, gy, gz, pi, i, t, slot, grp_shift, live, spos, cell_start, cell_end, n, h, h2, h2_lo,
                       h2_hi, nbr, n, acc, count);
    }
    acc += __shfl_xor(acc, 1, kWave);
    acc += __shfl_xor(acc, 2, kWave);
    acc += __shfl_xor(acc, 4, kWave);
    if (!live || slot != 0) return;
    float rho, prs;
    if (count > 2) {
        rho = fmaxf(1.0f, p.mass * p.kernelConstant * acc);
        prs = p.pressure + p.stiffness * (rho - p.density);
    } else {
        rho = p.density;
        prs = p.pressure;
    }
    spos[i].w = rho;
    svel[i].w = prs;
    ncount[t] = count;
    if (counts_out) counts_out[i] = count;
}

template <int ORDER, int CH>
__device__ inline void forces_of(uint32_t t, uint32_t i, const float4* __restrict__ spos,
                                 const float4* __restrict__ svel, const uint32_t* __restrict__ cell_start,
                                 const uint32_t* __restrict__ cell_end, const uint32_t* __restrict__ nbr,
                                 const uint32_t* __restrict__ ncount, uint32_t n, float h, const float4& pi,
                                 const float4& vi, float3& P, float3& V)
{
    const uint32_t count = ncount[t];
    if (count <= kNbrCap)
        sph_force_sum_nbr<CH>(spos, svel, nbr, n, count, t, pi, vi, h, P, V);
    else
        sph_force_sum(spos, svel, cell_start, cell_end, i, n, pi, vi, h, P, V);
}

template <int ORDER, int CH>
__global__ __launch_bounds__(kThreads) void k_forces_nbr(const float4* __restrict__ spos,
                                                         const float4* __restrict__ svel, float4* __restrict__ force,
                                                         const uint32_t* __restrict__ index,
                                                         const uint32_t* __restrict__ cell_start,
                                                         const uint32_t* __restrict__ cell_end,
                                                         const uint32_t* __restrict__ nbr,
                                                         const uint32_t* __restrict__ ncount, uint32_t n, mfl_params p,
                                                         const uint32_t* __restrict__ porder,
                                                         const uint32_t* __restrict__ inv)
{
    const uint32_t t = xcd_block() * kThreads + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = sorted_slot<ORDER>(t, n, porder, inv);
    if (i >= n) return;
    const float4 pi = spos[i], vi = svel[i];
    float3 P, V;
    forces_of<ORDER, CH>(t, i, spos, svel, cell_start, cell_end, nbr, ncount, n, p.cellSize, pi, vi, P, V);
    const uint32_t o = caller_slot<ORDER>(t, i, porder, index);
    if (o >= n) return;
    force[o] = add_sph_force(force[o], P, V, p);
}

// Forces + clip box + integrate force + gravity + integrate velocity for one particle: sorted slot i,
// caller-order slot o = index[i].
template <int ORDER, int CH>
__global__ __launch_bounds__(kThreads) void k_forces_step(const float4* __restrict__ spos,
                                                          const float4* __restrict__ svel, float4* __restrict__ pos,
                                                          float4* __restrict__ vel, float4* __restrict__ force,
                                                          const uint32_t* __restrict__ index,
                                                          const uint32_t* __restrict__ cell_start,
                                                          const uint32_t* __restrict__ cell_end,
                                                          const uint32_t* __restrict__ nbr,
                                                          const uint32_t* __restrict__ ncount, uint32_t n,
                                                          mfl_params p, float4 box_min, float4 box_max,
                                                          float4 gravity_force, const uint32_t* __restrict__ porder,
                                                          const uint32_t* __restrict__ inv)
{
    const uint32_t t = xcd_block() * kThreads + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = sorted_slot<ORDER>(t, n, porder, inv);
    if (i >= n) return;
    const float4 pi = spos[i], vi = svel[i];
    float3 P, V;
    forces_of<ORDER, CH>(t, i, spos, svel, cell_start, cell_end, nbr, ncount, n, p.cellSize, pi, vi, P, V);
    const uint32_t o = caller_slot<ORDER>(t, i, porder, index);
    if (o >= n) return;
    float4 f = add_sph_force(force[o], P, V, p);
    f = clip_box(f, pi, vi, p.particleRadius, box_min, box_max);
    const float4 v = integrate_force(vel[o], f, p);
    force[o] = gravity_force;
    vel[o] = v;
    pos[o] = f4_add(pos[o], f4_scale(v, p.deltaTime));
}

// Walks of a neighbourhood that overflowed the tile, from global memory. Kept out of line so that the
// rarely taken walk does not set the register budget of the tile kernels.
// This lane's kernel sum and its group's visit count (the same in every lane of the group); records the
// visits of traversal slot t in nbr when nbr is not null. Returned by value so that the callers' sums
// stay in registers.
struct VisitSums {
    float acc;
    uint32_t count;
};

__device__ __noinline__ VisitSums density_fallback(int cx, int cy, int cz, float4 pi, uint32_t i, uint32_t t,
                                                  uint32_t slot, uint32_t grp, const float4* __restrict__ spos,
                                                  const uint32_t* __restrict__ cell_start,
                                                  const uint32_t* __restrict__ cell_end, uint32_t n, float h,
                                                  uint32_t* __restrict__ nbr)
{
    const float h2 = h * h;
    VisitSums r{0.f, 0u};
    density_visits(cx, cy, cz, pi, i, t, slot, grp * kGroupLanes, true, spos, cell_start, cell_end, n, h, h2,
                   h2 * 0.9999996f, h2 * 1.0000005f, nbr, n, r.acc, r.count, nbr ? kNbrCap : 0u);
    return r;
}

struct PairSums {
    float3 P, V;
};

__device__ __noinline__ PairSums force_fallback(const float4* __restrict__ spos, const float4* __restrict__ svel,
                                               const uint32_t* __restrict__ cell_start,
                                               const uint32_t* __restrict__ cell_end, uint32_t i, uint32_t n,
                                               float4 pi, float4 vi, float h)
{
    PairSums r;
    sph_force_sum(spos, svel, cell_start, cell_end, i, n, pi, vi, h, r.P, r.V);
    return r;
}

// Particle of traversal slot t: its sorted slot i (n when t is past the end) and its position. The
// caller-order position pos[o] equals spos[i] (the reorder stage copies it), so it is read in parallel
// with the inverse permutation instead of after it.
template <int ORDER>
__device__ inline float4 slot_particle(uint32_t t, uint32_t n, const uint32_t* __restrict__ porder,
                                       const uint32_t* __restrict__ inv, const float4* __restrict__ pos,
                                       const float4* __restrict__ spos, uint32_t& i)
{
    i = n;
    if (t >= n) return make_float4(0.f, 0.f, 0.f, 0.f);
    if (ORDER == kOrderSorted) {
        i = t;
        return spos[t];
    }
    const uint32_t o = ORDER == kOrderCaller ? t : porder[t];
    if (o >= n) return make_float4(0.f, 0.f, 0.f, 0.f);
    i = inv[o];
    return pos[o];
}

// One lane walks particle i's 27 bucket ranges from global memory (the reference's loop, cl:161-190) and
// records its visits for traversal slot t: the overflow path of the 4-lane layout when visits are
// recorded (without a record the lanes split the walk, density_walk_split).
__device__ __noinline__ VisitSums density_walk(int cx, int cy, int cz, float4 pi, uint32_t i, uint32_t t,
                                              const float4* __restrict__ spos, const uint32_t* __restrict__ cell_start,
                                              const uint32_t* __restrict__ cell_end, uint32_t n, float h,
                                              uint32_t* __restrict__ nbr)
{
    const float h2 = h * h;
    const float h2_lo = h2 * 0.9999996f, h2_hi = h2 * 1.0000005f;
    VisitSums r{0.f, 0u};
    for (int k = 0; k < 27; k++) {
        const uint32_t b = cell_hash(cx + k % 3 - 1, cy + (k / 3) % 3 - 1, cz + k / 9 - 1);
        const uint32_t st = cell_start[b];
        if (st == kEmpty) continue;
        const uint32_t end = min(cell_end[b], n);
        for (uint32_t j = st; j < end; j++) {
            const float4 pj = spos[j];
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float r2 = dist2(dx, dy, dz);
            if (j != i && accept_r2(r2, h, h2_lo, h2_hi)) {
                const float w = h2 - r2;
                r.acc = fmaf(w * w, w, r.acc);
                if (nbr && r.count < kNbrCap) nbr[nbr_at(t, r.count, n)] = j;
                r.count++;
            }
        }
    }
    return r;
}

// Outputs of the tile density pass: none, recorded visit lists, or each cell's kept entries plus the
// lanes' acceptance bits (kSegCap, kMetaWalk, kMaxWindows in tile.hpp).
enum : int { kOutNone = 0, kOutList = 1, kOutKept = 2 };

// Density pass over LDS-staged neighbourhoods (cl:150-200). The wave is split into 64 / W segments of
// 8 particles (W / 8 lanes each); each segment stages its cell's neighbourhood in its own `depth + 8`
// float4 of dynamic LDS plus a StageAux. A neighbourhood that keeps more than `depth` entries is tested
// `depth` kept entries at a time (sweep_windows); one that needs more than kMaxWindows windows, or holds
// more than 512 visits, falls back to walking the ranges from global memory. LIST = true also records
// each particle's accepted visits (nbr, ncount by traversal slot) for the recorded-visit force pass.
template <int ORDER, int OUT, int W>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_density_tile(const float4* __restrict__ pos, float4* __restrict__ spos,
                                                           float4* __restrict__ svel,
                                                           const uint32_t* __restrict__ cell_start,
                                                           const uint32_t* __restrict__ cell_end, uint32_t n,
                                                           mfl_params p, uint32_t depth,
                                                           const uint32_t* __restrict__ porder,
                                                           const uint32_t* __restrict__ inv, uint32_t* __restrict__ nbr,
                                                           uint32_t* __restrict__ ncount, uint32_t* __restrict__ kept_out,
                                                           uint64_t* __restrict__ hitmask, uint64_t* __restrict__ hitx,
                                                           uint32_t* __restrict__ meta, float cut,
                                                           uint32_t* __restrict__ counts_out)
{
    constexpr bool LIST = OUT == kOutList;
    constexpr uint32_t P = W / 8;  // lanes per particle
    constexpr uint32_t SEGS = kWave / W;
    extern __shared__ float4 lds[];
    const Lanes<W> L(threadIdx.x & (kWave - 1));
    const uint32_t nseg = (blockDim.x / kWave) * SEGS, sid = (threadIdx.x / kWave) * SEGS + L.seg;
    float4* tile = lds + (size_t)sid * (depth + kGroupLanes);
    StageAux* aux = reinterpret_cast<StageAux*>(lds + (size_t)nseg * (depth + kGroupLanes)) + sid;
    const uint32_t grp = L.ls / P, slot = L.ls % P;
    const uint32_t wave = (xcd_block() * blockDim.x + threadIdx.x) / kWave;
    const uint32_t t = wave * (kWave / P) + L.seg * 8 + grp;
    uint32_t i;
    const float4 pi = slot_particle<ORDER>(t, n, porder, inv, pos, spos, i);
    const bool live = i < n;
    const float h = p.cellSize;
    const float h2 = h * h;
    const int gx = grid_coord(pi.x, h), gy = grid_coord(pi.y, h), gz = grid_coord(pi.z, h);
    float acc = 0.f;
    uint32_t count = 0;  // this lane's accepted visits (summed over the particle's lanes at the end)
    uint32_t used = 0;   // OUT == kOutKept: entries of this segment's kept buffer already written
    const uint32_t gseg = wave * SEGS + L.seg;
    uint64_t todo = L.bits(__ballot(live));
    while (__ballot(todo != 0) != 0) {
        const SegCell c = next_cell_seg<W>(L, todo, live, gx, gy, gz, pi);
        todo &= ~L.bits(__ballot(c.member));
        const uint32_t total = stage_seg_ranges<W>(L, c, cell_start, cell_end, n, aux);
        const uint32_t kept = total == kStageOverflow ? kStageOverflow
                                                      : stage_seg_fill<W>(L, total, c.lo, c.hi, h2 * kKeepMargin, spos, 0u, depth, tile, aux);
        if constexpr (OUT == kOutKept) {
            // the cell's kept entries go after the segment's earlier cells when they all fit
            if (kept != kStageOverflow && used + kept > kSegCap && c.member && slot == 0) meta[t] = kMetaWalk;
        }
        uint32_t at_base = 0;  // LIST: the particle's visits recorded in earlier windows
        const bool fits = sweep_windows<W>(
            L, kept, h2 * kKeepMargin, spos, depth, tile, aux, pi, i, h2, cut,
            [&](uint32_t w, uint32_t nw, uint32_t trips, uint64_t hits, float a) {
                if (c.member) {
                    acc += a;
                    count += (uint32_t)__popcll(hits);
                }
                if constexpr (OUT == kOutKept) {
                    // the window's kept entries once for the segment, the acceptance bits per lane
                    if (used + kept <= kSegCap) {
                        for (uint32_t q = L.ls; q < nw; q += W)
                            kept_out[(size_t)gseg * kSegCap + used + w * depth + q] = __float_as_uint(tile[q].w);
                        if (c.member) {
                            if (w == 0) hitmask[(size_t)t * P + slot] = hits;
                            else hitx[hitx_at(t, P, slot, w)] = hits;
                        }
                    }
                }
                if (LIST && c.member) {
                    // this lane's accepted visits go after those of the particle's earlier windows and lower lanes
                    const uint32_t mine = (uint32_t)__popcll(hits);
                    uint32_t at = mine;
#pragma unroll
                    for (uint32_t o = 1; o < P; o <<= 1) {
                        const uint32_t up = (uint32_t)__shfl_up((int)at, o, P);
                        if (slot >= o) at += up;
                    }
                    const uint32_t all = (uint32_t)__shfl((int)at, (int)(L.lane | (P - 1)), kWave);  // the particle's total
                    at += at_base - mine;
                    at_base += all;
                    while (hits) {
                        const uint32_t m = trips - 1u - (uint32_t)__builtin_ctzll(hits);
                        hits &= hits - 1ull;
                        if (at < kNbrCap) nbr[nbr_at(t, at, n)] = __float_as_uint(tile[m * P + slot].w);
                        at++;
                    }
                }
            });
        if constexpr (OUT == kOutKept) {
            if (fits && used + kept <= kSegCap) {
                if (c.member && slot == 0) meta[t] = used | (kept << 16);
                used += kept;
            }
        }
        if (!fits && c.member) {
            if (OUT == kOutKept && slot == 0) meta[t] = kMetaWalk;
            if constexpr (W == kWave) {
                const VisitSums f = density_fallback(c.cx, c.cy, c.cz, pi, i, t, slot, grp, spos, cell_start, cell_end,
                                                     n, h, LIST ? nbr : nullptr);
                acc += f.acc;
                count += slot == 0 ? f.count : 0u;  // every lane of the group holds the group's count
            } else if constexpr (LIST) {
                if (slot == 0) {
                    const VisitSums f = density_walk(c.cx, c.cy, c.cz, pi, i, t, spos, cell_start, cell_end, n, h, nbr);
                    acc += f.acc;
                    count += f.count;
                }
            } else {  // no visits to record: the particle's P lanes split the walk
                const WalkSums f = density_walk_split(c.cx, c.cy, c.cz, pi, i, spos, cell_start, cell_end, n, h, slot, P);
                acc += f.acc;
                count += f.count;
            }
        }
        wave_sync();  // the next cell's neighbourhood overwrites the tile
    }
    acc = particle_sum<W>(acc);
    count = particle_sum_u<W>(count);
    if (!live || slot != 0) return;
    float rho, prs;
    if (count > 2) {
        rho = fmaxf(1.0f, p.mass * p.kernelConstant * acc);
        prs = p.pressure + p.stiffness * (rho - p.density);
    } else {
        rho = p.density;
        prs = p.pressure;
    }
    spos[i].w = rho;
    svel[i].w = prs;
    if (LIST) ncount[t] = count;
    if (counts_out) counts_out[i] = count;
}

// Force pass over LDS-staged neighbourhoods (cl:202-255). STEP = true fuses clip box, integrate force,
// gravity and integrate velocity (cl:258-361), as k_forces_step does. Each wave owns `depth` entries of
// 36 bytes (position, velocity, index).
template <int ORDER, bool STEP>
__global__ __launch_bounds__(kThreads) void k_forces_tile(const float4* __restrict__ spos,
                                                          const float4* __restrict__ svel, float4* __restrict__ pos,
                                                          float4* __restrict__ vel, float4* __restrict__ force,
                                                          const uint32_t* __restrict__ index,
                                                          const uint32_t* __restrict__ cell_start,
                                                          const uint32_t* __restrict__ cell_end, uint32_t n,
                                                          mfl_params p, uint32_t depth, float4 box_min, float4 box_max,
                                                          float4 gravity_force, const uint32_t* __restrict__ porder,
                                                          const uint32_t* __restrict__ inv)
{
    extern __shared__ float4 lds[];
    const uint32_t waves = blockDim.x / kWave, wv = threadIdx.x / kWave;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    float4* tpos = lds + (size_t)wv * 2 * depth;
    float4* tvel = tpos + depth;
    uint32_t* tj = reinterpret_cast<uint32_t*>(lds + (size_t)waves * 2 * depth) + (size_t)wv * depth;
    StageAux* aux = reinterpret_cast<StageAux*>(reinterpret_cast<uint32_t*>(lds + (size_t)waves * 2 * depth) +
                                                (size_t)waves * depth) + wv;
    const uint32_t grp = lane / kGroupLanes, slot = lane % kGroupLanes;
    const uint32_t wave = (xcd_block() * blockDim.x + threadIdx.x) / kWave;
    const uint32_t t = wave * (kWave / kGroupLanes) + grp;
    uint32_t i;
    const float4 px = slot_particle<ORDER>(t, n, porder, inv, pos, spos, i);
    const bool live = i < n;
    // density and pressure of the particle itself (the w lanes), needed only after staging
    const float4 pi = live ? spos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 vi = live ? svel[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float h = p.cellSize;
    const float h2 = h * h;
    const float h2_lo = h2 * 0.9999996f, h2_hi = h2 * 1.0000005f;
    const int gx = grid_coord(px.x, h), gy = grid_coord(px.y, h), gz = grid_coord(px.z, h);
    float3 P = make_float3(0.f, 0.f, 0.f), V = make_float3(0.f, 0.f, 0.f);
    uint64_t todo = __ballot(live);
    while (todo) {
        const CellGroup c = next_cell(todo, live, gx, gy, gz, px);
        todo &= ~__ballot(c.member);
        const uint32_t kept = stage_neighbourhood<true>(c.cx, c.cy, c.cz, c.lo, c.hi, h2 * kKeepMargin, spos, svel,
                                                        cell_start, cell_end, n, lane, depth, tpos, tvel, tj, aux);
        if (c.member) {
            if (kept <= depth) {
                for (uint32_t q = slot; q < kept; q += kGroupLanes) {
                    const float4 pj = tpos[q];
                    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    const float r2 = dist2(dx, dy, dz);
                    if (tj[q] != i && accept_r2(r2, h, h2_lo, h2_hi))
                        pair_force(pi, vi, pj, tvel[q], dx, dy, dz, r2, h, P, V);
                }
            } else if (slot == 0) {
                const PairSums f = force_fallback(spos, svel, cell_start, cell_end, i, n, pi, vi, h);
                P = f.P;
                V = f.V;
            }
        }
        wave_sync();
    }
    P = make_float3(group_sum(P.x), group_sum(P.y), group_sum(P.z));
    V = make_float3(group_sum(V.x), group_sum(V.y), group_sum(V.z));
    if (!live || slot != 0) return;
    const uint32_t o = caller_slot<ORDER>(t, i, porder, index);
    if (o >= n) return;
    float4 f = add_sph_force(force[o], P, V, p);
    if (!STEP) {
        force[o] = f;
        return;
    }
    f = clip_box(f, pi, vi, p.particleRadius, box_min, box_max);
    const float4 v = integrate_force(vel[o], f, p);
    force[o] = gravity_force;
    vel[o] = v;
    pos[o] = f4_add(pos[o], f4_scale(v, p.deltaTime));
}

// Force pass over the kept entries of the tile density pass (cl:202-255, STEP: + cl:258-361): the same
// waves and segments as the density pass; each segment stages the positions (with density) and
// velocities (with pressure) of its cell's kept entries in LDS from the recorded kept list, and every
// lane sums the pair terms of the entries its acceptance bits mark (the density pass tested exactly
// these). Particles whose cell did not fit walk the 27 ranges.
template <int ORDER, int W, bool STEP>
__global__ __launch_bounds__(kThreads) void k_forces_kept(const float4* __restrict__ spos,
                                                          const float4* __restrict__ svel, float4* __restrict__ pos,
                                                          float4* __restrict__ vel, float4* __restrict__ force,
                                                          const uint32_t* __restrict__ index,
                                                          const uint32_t* __restrict__ cell_start,
                                                          const uint32_t* __restrict__ cell_end, uint32_t n,
                                                          mfl_params p, uint32_t depth, float4 box_min, float4 box_max,
                                                          float4 gravity_force, const uint32_t* __restrict__ porder,
                                                          const uint32_t* __restrict__ inv,
                                                          const uint32_t* __restrict__ kept_in,
                                                          const uint64_t* __restrict__ hitmask,
                                                          const uint64_t* __restrict__ hitx,
                                                          const uint32_t* __restrict__ meta)
{
    constexpr uint32_t P = W / 8;
    constexpr uint32_t SEGS = kWave / W;
    extern __shared__ float4 lds[];
    const Lanes<W> L(threadIdx.x & (kWave - 1));
    constexpr uint32_t chunk = kept_chunk<W>();  // small LDS per segment keeps more waves resident
    // the wave's tiles: positions of its segments' chunks, then their velocities
    float4* wpos = lds + (size_t)(threadIdx.x / kWave) * SEGS * 2 * chunk;
    float4* wvel = wpos + SEGS * chunk;
    float4* tpos = wpos + L.seg * chunk;
    float4* tvel = wvel + L.seg * chunk;
    const uint32_t grp = L.ls / P, slot = L.ls % P;
    const uint32_t wave = (xcd_block() * blockDim.x + threadIdx.x) / kWave;
    const uint32_t t = wave * (kWave / P) + L.seg * 8 + grp;
    const uint32_t gseg = wave * SEGS + L.seg;
    const uint32_t i = t < n ? sorted_slot<ORDER>(t, n, porder, inv) : n;
    const bool live = i < n;
    const float h = p.cellSize;
    const float4 pi = live ? spos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 vi = live ? svel[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t mt = live ? meta[t] : kMetaWalk;
    const bool staged = live && mt != kMetaWalk;
    const uint64_t hm = staged ? hitmask[(size_t)t * P + slot] : 0ull;
    float3 Pa = make_float3(0.f, 0.f, 0.f), Va = make_float3(0.f, 0.f, 0.f);
    uint64_t todo = L.bits(__ballot(staged));
    while (__ballot(todo != 0) != 0) {
        const bool active = todo != 0;
        const uint32_t src = active ? (uint32_t)(__ffsll((unsigned long long)todo) - 1) : 0u;
        const uint32_t gm = L.from(mt, src);
        const bool member = active && staged && mt == gm;
        todo &= ~L.bits(__ballot(member));
        const uint32_t off = active ? (gm & 0xFFFFu) : 0u, kept = active ? (gm >> 16) : 0u;
        const uint32_t kmax = seg_wave_max<W>(kept);
        // the density pass's windows of `depth` entries, each with its own word of acceptance bits
        for (uint32_t w0 = 0, w = 0; w0 < kmax; w0 += depth, w++) {
            const uint32_t nw = kept > w0 ? min(kept - w0, depth) : 0u, nmax = min(kmax - w0, depth);
            const uint32_t trips = (nw + P - 1) / P;
            const uint64_t hw = w == 0 ? hm : (member && nw ? hitx[hitx_at(t, P, slot, w)] : 0ull);
            const size_t kbase = (size_t)gseg * kSegCap + off + w0;
            // `chunk` entries at a time; the next chunk's index loads are in flight during a chunk's sweep
            uint32_t kidx[SEGS];
            kept_chunk_index<W>(kept_in, kbase, nw, 0u, L.lane, kidx);
            for (uint32_t c0 = 0; c0 < nmax; c0 += chunk) {
                kept_chunk_load<W>(kidx, nw, c0, L.lane, spos, svel, wpos, wvel);
                __builtin_amdgcn_s_waitcnt(0x0F70);  // the states are in LDS (vmcnt(0))
                // issued after that wait, so that no LDS read of the sweep waits for them
                if (c0 + chunk < nmax) kept_chunk_index<W>(kept_in, kbase, nw, c0 + chunk, L.lane, kidx);
                wave_sync();
                if (member) {
                    uint32_t bits = chunk_bits<P, chunk>(hw, trips, c0);
                    const float4* rp = tpos + slot;  // this lane's entries: k P + slot of the chunk
                    const float4* rv = tvel + slot;
                    while (bits) {
                        const uint32_t k = (uint32_t)__builtin_clz(bits);  // earliest trip first
                        bits ^= 0x80000000u >> k;
                        const float4 pj = rp[k * P], vj = rv[k * P];
                        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                        pair_force(pi, vi, pj, vj, dx, dy, dz, dist2(dx, dy, dz), h, Pa, Va);
                    }
                }
                wave_sync();  // the next chunk overwrites the tile
            }
        }
    }
    if (live && !staged) {  // the particle's P lanes split the walk of a cell that overflowed its tile
        const WalkForces f = force_walk_split(spos, svel, cell_start, cell_end, i, n, pi, vi, h, slot, P);
        Pa = f.P;
        Va = f.V;
    }
    Pa = make_float3(particle_sum<W>(Pa.x), particle_sum<W>(Pa.y), particle_sum<W>(Pa.z));
    Va = make_float3(particle_sum<W>(Va.x), particle_sum<W>(Va.y), particle_sum<W>(Va.z));
    if (!live || slot != 0) return;
    const uint32_t o = caller_slot<ORDER>(t, i, porder, index);
    if (o >= n) return;
    float4 f = add_sph_force(force[o], Pa, Va, p);
    if (!STEP) {
        force[o] = f;
        return;
    }
    f = clip_box(f, pi, vi, p.particleRadius, box_min, box_max);
    const float4 v = integrate_force(vel[o], f, p);
    force[o] = gravity_force;
    vel[o] = v;
    pos[o] = f4_add(pos[o], f4_scale(v, p.deltaTime));
}

// Morton code of each particle's cell, coordinates taken modulo 2^mbits (mbits <= 10).
__global__ __launch_bounds__(kThreads) void k_morton_keys(const float4* __restrict__ pos, uint32_t n, float h,
                                                          uint32_t mbits, uint32_t* __restrict__ key)
{
    const uint32_t o = blockIdx.x * kThreads + threadIdx.x;
    if (o >= n) return;
    const float4 q = pos[o];
    const uint32_t m = (1u << mbits) - 1u;
    key[o] = spread_bits3((uint32_t)grid_coord(q.x, h) & m) | (spread_bits3((uint32_t)grid_coord(q.y, h) & m) << 1) |
             (spread_bits3((uint32_t)grid_coord(q.z, h) & m) << 2);
}

__global__ __launch_bounds__(kThreads) void k_clip_box(const float4* __restrict__ spos, const float4* __restrict__ svel,
                                                       float4* __restrict__ force, const uint32_t* __restrict__ index,
                                                       uint32_t n, mfl_params p, float4 box_min, float4 box_max)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = index[i];
    if (o >= n) return;
    force[o] = clip_box(force[o], spos[i], svel[i], p.particleRadius, box_min, box_max);
}

__global__ __launch_bounds__(kThreads) void k_integrate_force(float4* __restrict__ vel, const float4* __restrict__ force,
                                                              uint32_t n, mfl_params p)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    vel[i] = integrate_force(vel[i], force[i], p);
}

__global__ __launch_bounds__(kThreads) void k_gravity(float4* __restrict__ force, uint32_t n, float4 gravity_force)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) force[i] = gravity_force;
}

__global__ __launch_bounds__(kThreads) void k_integrate_velocity(float4* __restrict__ pos,
                                                                 const float4* __restrict__ vel, uint32_t n,
                                                                 mfl_params p)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    pos[i] = f4_add(pos[i], f4_scale(vel[i], p.deltaTime));
}

}  // namespace

// ---------------------------------------------------------------------------------------------------

StageTimer::~StageTimer()
{
    drain();
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
    for (hipEvent_t e : open_)
        if (e) (void)hipEventDestroy(e);
}

hipEvent_t StageTimer::take()
{
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

bool StageTimer::begin(int stage, hipStream_t s)
{
    if (!(mask & (1u << stage))) return false;
    hipEvent_t e = take();
    if (!e) return false;
    (void)hipEventRecord(e, s);
    open_[stage] = e;
    return true;
}

void StageTimer::end(int stage, hipStream_t s)
{
    hipEvent_t a = open_[stage];
    if (!a) return;
    open_[stage] = nullptr;
    hipEvent_t b = take();
    if (!b) {
        pool_.push_back(a);
        return;
    }
    (void)hipEventRecord(b, s);
    pending_.push_back({stage, a, b});
}

bool StageTimer::drain()
{
    bool ok = true;
    for (const Pending& p : pending_) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ms_[p.stage] += ms;
            count_[p.stage] += 1;
        } else {
            ok = false;
        }
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.clear();
    return ok;
}

int StageTimer::read(int stage, double* ms, uint64_t* launches, bool reset)
{
    if (stage < 0 || stage >= MFL_STAGE_COUNT) {
        set_error("unknown stage " + std::to_string(stage));
        return 0;
    }
    if (!drain()) {
        set_error("stage timing events failed");
        return 0;
    }
    if (ms) *ms = ms_[stage];
    if (launches) *launches = count_[stage];
    if (reset) {
        for (int i = 0; i < MFL_STAGE_COUNT; i++) {
            ms_[i] = 0.0;
            count_[i] = 0;
        }
    }
    return 1;
}

Fluid3D::Fluid3D()
{
    mfl_params_default(&params);
    if (!params_buf_.allocate(64)) return;
    allocate_internal();
}

Fluid3D::~Fluid3D()
{
    drop_graph();
}

bool Fluid3D::allocate_internal()
{
    drop_graph();
    const size_t v4 = n_ * 16, u1 = n_ * 4;
    if (!cell_.allocate(u1) || !index_.allocate(u1) || !spos_.allocate(v4) || !svel_.allocate(v4)) return false;
    if (!cell_start_.allocate(kBuckets * 4) || !cell_end_.allocate(kBuckets * 4)) return false;
    if (!nbr_.allocate((size_t)kNbrCap * u1) || !ncount_.allocate(u1)) return false;
    // kept entries of the tile density pass: kSegCap per 8-particle segment, 64-bit bits per lane
    // and kMaxWindows - 1 more words per lane for neighbourhoods tested in several windows
    // (segments of every wave of up to 64 particles, so that any segment's region exists)
    if (!kept_.allocate((n_ + 63) / 64 * 8 * kSegCap * 4) || !hitmask_.allocate(u1 / 4 * kGroupLanes * 8) ||
        !hitx_.allocate(u1 / 4 * kGroupLanes * (kMaxWindows - 1) * 8) || !meta_.allocate(u1))
        return false;
    if (!inv_.allocate(u1) || !porder_.allocate(u1) || !mkey_.allocate(u1) || !counts_.allocate(u1)) return false;
    if (n_ <= kSelectMax) {  // the selecting run pass reads kSelectMax keys: those past n never match
        if (!skey_.allocate(kSelectMax * 4) || !hip_ok(hipMemset(skey_.ptr, 0xFF, kSelectMax * 4), "hipMemset"))
            return false;
        tkey_.release();
        tval_.release();
        tcount_.release();
        tstart_.release();
    } else {  // the hash kernel's tile-ordered keys and indices, and per tile and high digit counts / starts
        skey_.release();
        const size_t max_tiles = (n_ + kThreads * 4 - 1) / (kThreads * 4);  // the smallest tiles
        if (!tkey_.allocate(u1) || !tval_.allocate(u1) || !tcount_.allocate(max_tiles * 512 * 4) ||
            !tstart_.allocate(max_tiles * 512 * 4))
            return false;
    }
    porder_valid_ = false;
    steps_since_refresh_ = 0;
    // the reference leaves the tables uninitialised (cpp:53-54); here they start empty
    if (!hip_ok(hipMemset(cell_start_.ptr, 0xFF, kBuckets * 4), "hipMemset") ||
        !hip_ok(hipMemset(cell_end_.ptr, 0xFF, kBuckets * 4), "hipMemset"))
        return false;
    for (DeviceBuffer* b : {&pos_, &vel_, &force_}) {
        if (!b->owned) continue;
        if (!b->allocate(v4)) return false;
        if (v4 && !hip_ok(hipMemset(b->ptr, 0, v4), "hipMemset")) return false;
    }
    if (!hip_ok(hipMemset(spos_.ptr, 0, v4), "hipMemset") || !hip_ok(hipMemset(svel_.ptr, 0, v4), "hipMemset"))
        return false;
    return hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

int Fluid3D::compile()
{
    if (!sort.compile()) return 0;
    mfl_params_default(&params);  // clInitFluid (cl:130-148)
    MFL_HIP(hipMemcpy(params_buf_.ptr, &params, sizeof(mfl_params), hipMemcpyHostToDevice));
    return 1;
}

int Fluid3D::set_particle_count(size_t n)
{
    if (n > 0xFFFFFFFFull / 16) {
        set_error("particle count too large");
        return 0;
    }
    n_ = n;
    // up to 2^17 particles (32 bytes of position and velocity each: one XCD's L2) the hash-sorted order
    // loses no locality and, unlike a Morton order, never goes stale between refreshes
    if (!traversal_chosen) traversal = n <= ((size_t)1 << 17) ? MFL_TRAVERSAL_SORTED : MFL_TRAVERSAL_MORTON;
    return allocate_internal() ? 1 : 0;
}

DeviceBuffer* Fluid3D::select(int kind)
{
    switch (kind) {
    case MFL_POSITION: return &pos_;
    case MFL_VELOCITY: return &vel_;
    case MFL_FORCE: return &force_;
    case MFL_SORTED_POSITION: return &spos_;
    case MFL_SORTED_VELOCITY: return &svel_;
    case MFL_INDEX: return &index_;
    case MFL_CELL: return &cell_;
    case MFL_CELL_START: return &cell_start_;
    case MFL_CELL_END: return &cell_end_;
    case MFL_PARAMS: return &params_buf_;
    case MFL_NEIGHBOUR_COUNT: return &counts_;
    default: set_error("unknown buffer kind " + std::to_string(kind)); return nullptr;
    }
}

int Fluid3D::set_buffer(int kind, void* ptr, size_t bytes)
{
    if (kind != MFL_POSITION && kind != MFL_VELOCITY && kind != MFL_FORCE) {
        set_error("only the position, velocity and force buffers can be replaced");
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
    DeviceBuffer* b = select(kind);
    drop_graph();
    b->release();  // frees a program-owned buffer only (deleteBuffer, libCL/phys/oclFluid3D.h:106-112)
    b->ptr = ptr;
    b->bytes = bytes;
    b->owned = false;
    return 1;
}

int Fluid3D::get_buffer(int kind, void** ptr, size_t* bytes)
{
    DeviceBuffer* b = select(kind);
    if (!b) return 0;
    if (ptr) *ptr = b->ptr;
    if (bytes) *bytes = kind == MFL_PARAMS ? sizeof(mfl_params) : b->bytes;
    return 1;
}

int Fluid3D::write_buffer(int kind, const void* host, size_t bytes)
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

int Fluid3D::read_buffer(int kind, void* host, size_t bytes)
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

bool Fluid3D::check_user_buffers()
{
    if (n_ == 0 || n_ % kGroup != 0) {
        set_error("particle count " + std::to_string(n_) + " is not a non-zero multiple of 256 (cLocalSize)");
        return false;
    }
    for (DeviceBuffer* b : {&pos_, &vel_, &force_}) {
        if (!b->ptr || b->bytes < n_ * 16) {
            set_error("a position/velocity/force buffer is smaller than the particle count");
            return false;
        }
    }
    return true;
}

// Particles per thread of the hash kernel's tile (k_hash_tiles): the sort's tile-size knob, else by size.
int Fluid3D::tile_items(size_t n) const
{
    const int r = sort.tuning.items;
    if (r == 4 || r == 8 || r == 16) return r;
    return n >= ((size_t)8 << 20) ? 16 : n >= ((size_t)1 << 20) ? 8 : 4;
}

int Fluid3D::enqueue_front(hipStream_t s)
{
    const uint32_t n = (uint32_t)n_;
    const dim3 g = grid_for(n), b(kThreads);
    auto* pos = pos_.as<float4>();
    auto* vel = vel_.as<float4>();
    auto* spos = spos_.as<float4>();
    auto* svel = svel_.as<float4>();
    auto* cell = cell_.as<uint32_t>();
    auto* index = index_.as<uint32_t>();
    auto* cs = cell_start_.as<uint32_t>();
    auto* ce = cell_end_.as<uint32_t>();
    const uint32_t clear = clear_count < kBuckets ? clear_count : kBuckets;

    // The keys (< 2^18) are ordered by their high 9 bits tile by tile inside the hash kernel, then
    // k_bucket_runs assembles each high digit's run from the tiles' segments, orders it by the low 9 bits
    // and writes the bucket bounds and the inverse permutation from the same histograms. Small steps
    // keep the keys in caller order and every run block selects its own.
    const bool select = n <= kSelectMax;
    const int items = tile_items(n);
    const uint32_t tile_keys = kThreads * (uint32_t)items, ntiles = (n + tile_keys - 1) / tile_keys;
    auto* tkey = tkey_.as<uint32_t>();
    auto* tval = tval_.as<uint32_t>();
    auto* tcnt = tcount_.as<uint32_t>();
    auto* tstart = tstart_.as<uint32_t>();
    auto* skey = skey_.as<uint32_t>();
    timer.begin(MFL_STAGE_HASH, s);
    if (select) {
        hipLaunchKernelGGL(k_hash_keys<4>, dim3((n + kThreads * 4 - 1) / (kThreads * 4)), b, 0, s, pos, n,
                           params.cellSize, skey, params_buf_.as<mfl_params>(), params, cs, ce, clear);
    } else {
#define MFL_HASH(IT)                                                                                              \
    hipLaunchKernelGGL(k_hash_tiles<IT>, dim3(ntiles), b, 0, s, pos, n, params.cellSize, params_buf_.as<mfl_params>(), \
                       params, cs, ce, clear, tkey, tval, tcnt, tstart, ntiles)
        switch (items) {
        case 16: MFL_HASH(16); break;
        case 8: MFL_HASH(8); break;
        default: MFL_HASH(4); break;
        }
#undef MFL_HASH
    }
    timer.end(MFL_STAGE_HASH, s);
    MFL_HIP(hipGetLastError());
    // the index buffer starts as the identity (cl:63) and is generated by the run pass
    timer.begin(MFL_STAGE_SORT, s);
    const bool gather = n < kScatterCopyMin;
    auto* inv = inv_.as<uint32_t>();
    if (select)
        hipLaunchKernelGGL((k_bucket_runs<true, kRunsSelect>), dim3(kRunDigits), dim3(kRunThreads), 0, s, skey,
                           nullptr, nullptr, nullptr, 0u, 0u, cell, index, cs, ce, inv, pos, vel, spos, svel);
    else if (gather)
        hipLaunchKernelGGL((k_bucket_runs<true, kRunsTiles>), dim3(kRunDigits), dim3(kRunThreads), 0, s, tkey, tval,
                           tcnt, tstart, ntiles, tile_keys, cell, index, cs, ce, inv, pos, vel, spos, svel);
    else
        hipLaunchKernelGGL((k_bucket_runs<false, kRunsTiles>), dim3(kRunDigits), dim3(kRunThreads), 0, s, tkey, tval,
                           tcnt, tstart, ntiles, tile_keys, cell, index, cs, ce, inv, pos, vel, spos, svel);
    MFL_HIP(hipGetLastError());
    timer.end(MFL_STAGE_SORT, s);
    // bucket bounds come out of the sort (the BOUNDS stage stays empty)
    timer.begin(MFL_STAGE_REORDER, s);
    if (!gather) hipLaunchKernelGGL(k_scatter_sorted, g, b, 0, s, inv, pos, vel, spos, svel, n);
    timer.end(MFL_STAGE_REORDER, s);
    MFL_HIP(hipGetLastError());
    return 1;
}

int Fluid3D::enqueue_step(hipStream_t s)
{
    const uint32_t n = (uint32_t)n_;
    const dim3 g = grid_for(n), b(kThreads);
    const float4 mn = to_f4(box_min), mx = to_f4(box_max), gf = gravity_force_of(gravity);
    auto* pos = pos_.as<float4>();
    auto* vel = vel_.as<float4>();
    auto* force = force_.as<float4>();
    auto* spos = spos_.as<float4>();
    auto* svel = svel_.as<float4>();
    auto* index = index_.as<uint32_t>();
    auto* cs = cell_start_.as<uint32_t>();
    auto* ce = cell_end_.as<uint32_t>();
    if (!enqueue_front(s)) return 0;
    const dim3 gx = grid_xcd(n);
    uint32_t* counts_out = record_counts ? counts_.as<uint32_t>() : nullptr;
    auto* nbr = nbr_.as<uint32_t>();
    auto* ncount = ncount_.as<uint32_t>();
    auto* inv = inv_.as<uint32_t>();
    auto* porder = porder_.as<uint32_t>();
    const int order = (traversal == MFL_TRAVERSAL_MORTON && !porder_valid_) ? MFL_TRAVERSAL_CALLER : traversal;
    if (neighbour_pass != MFL_PASS_LIST) return enqueue_tile_passes(s, order);
    timer.begin(MFL_STAGE_DENSITY, s);
    if (density_lanes == 1) {
        switch (order) {
        case MFL_TRAVERSAL_SORTED:
            hipLaunchKernelGGL(k_density_nbr<kOrderSorted>, gx, b, 0, s, spos, svel, cs, ce, n, params, nbr, ncount,
                               porder, inv, counts_out);
            break;
        case MFL_TRAVERSAL_CALLER:
            hipLaunchKernelGGL(k_density_nbr<kOrderCaller>, gx, b, 0, s, spos, svel, cs, ce, n, params, nbr, ncount,
                               porder, inv, counts_out);
            break;
        default:
            hipLaunchKernelGGL(k_density_nbr<kOrderListed>, gx, b, 0, s, spos, svel, cs, ce, n, params, nbr, ncount,
                               porder, inv, counts_out);
        }
    } else {
        const dim3 gg = grid_xcd((size_t)n * kGroupLanes);
        switch (order) {
        case MFL_TRAVERSAL_SORTED:
            hipLaunchKernelGGL(k_density_grp<kOrderSorted>, gg, b, 0, s, spos, svel, cs, ce, n, params, nbr, ncount,
                               porder, inv, counts_out);
            break;
        case MFL_TRAVERSAL_CALLER:
            hipLaunchKernelGGL(k_density_grp<kOrderCaller>, gg, b, 0, s, spos, svel, cs, ce, n, params, nbr, ncount,
                               porder, inv, counts_out);
            break;
        default:
            hipLaunchKernelGGL(k_density_grp<kOrderListed>, gg, b, 0, s, spos, svel, cs, ce, n, params, nbr, ncount,
                               porder, inv, counts_out);
        }
    }
    timer.end(MFL_STAGE_DENSITY, s);
    return enqueue_list_forces(s, order);
}

int Fluid3D::enqueue_list_forces(hipStream_t s, int order)
{
    const uint32_t n = (uint32_t)n_;
    const dim3 gx = grid_xcd(n), b(kThreads);
    const float4 mn = to_f4(box_min), mx = to_f4(box_max), gf = gravity_force_of(gravity);
    auto* pos = pos_.as<float4>();
    auto* vel = vel_.as<float4>();
    auto* force = force_.as<float4>();
    auto* spos = spos_.as<float4>();
    auto* svel = svel_.as<float4>();
    auto* index = index_.as<uint32_t>();
    auto* cs = cell_start_.as<uint32_t>();
    auto* ce = cell_end_.as<uint32_t>();
    auto* nbr = nbr_.as<uint32_t>();
    auto* ncount = ncount_.as<uint32_t>();
    auto* inv = inv_.as<uint32_t>();
    auto* porder = porder_.as<uint32_t>();
    timer.begin(MFL_STAGE_FORCES, s);
    if (integrate_cb) {
#define MFL_FORCES_NBR(O)                                                                                     \
    if (force_chunks == 2)                                                                                    \
        hipLaunchKernelGGL((k_forces_nbr<O, 2>), gx, b, 0, s, spos, svel, force, index, cs, ce, nbr, ncount, n,  \
                           params, porder, inv);                                                              \
    else                                                                                                      \
        hipLaunchKernelGGL((k_forces_nbr<O, 1>), gx, b, 0, s, spos, svel, force, index, cs, ce, nbr, ncount, n,  \
                           params, porder, inv)
        switch (order) {
        case MFL_TRAVERSAL_SORTED: MFL_FORCES_NBR(kOrderSorted); break;
        case MFL_TRAVERSAL_CALLER: MFL_FORCES_NBR(kOrderCaller); break;
        default: MFL_FORCES_NBR(kOrderListed);
        }
#undef MFL_FORCES_NBR
        timer.end(MFL_STAGE_FORCES, s);
        MFL_HIP(hipGetLastError());
        (void)integrate_cb(handle, (void*)s, integrate_user);  // return value ignored, cpp:325-328
        return 1;
    }
#define MFL_FORCES_STEP(O)                                                                                    \
    if (force_chunks == 2)                                                                                    \
        hipLaunchKernelGGL((k_forces_step<O, 2>), gx, b, 0, s, spos, svel, pos, vel, force, index, cs, ce, nbr,  \
                           ncount, n, params, mn, mx, gf, porder, inv);                                       \
    else                                                                                                      \
        hipLaunchKernelGGL((k_forces_step<O, 1>), gx, b, 0, s, spos, svel, pos, vel, force, index, cs, ce, nbr,  \
                           ncount, n, params, mn, mx, gf, porder, inv)
    switch (order) {
    case MFL_TRAVERSAL_SORTED: MFL_FORCES_STEP(kOrderSorted); break;
    case MFL_TRAVERSAL_CALLER: MFL_FORCES_STEP(kOrderCaller); break;
    default: MFL_FORCES_STEP(kOrderListed);
    }
#undef MFL_FORCES_STEP
    timer.end(MFL_STAGE_FORCES, s);
    MFL_HIP(hipGetLastError());
    return 1;
}

// Tile passes: `block_waves` waves per workgroup, grid a multiple of 8 blocks. The density pass holds
// 64 / lanes particles per wave in 8-particle segments (lanes = 8: one segment, lanes = 4: two); the
// force pass holds 8 particles per wave.
static size_t tile_lds_bytes(bool forces, uint32_t waves, uint32_t depth, uint32_t lanes = 8)
{
    if (forces) return (size_t)waves * (depth * (2 * sizeof(float4) + sizeof(uint32_t)) + sizeof(StageAux));
    const size_t segs = (size_t)waves * (8 / lanes);
    return segs * ((depth + kGroupLanes) * sizeof(float4) + sizeof(StageAux));  // + 8 end markers
}

static dim3 tile_grid(size_t n, uint32_t particles_per_wave, uint32_t block_waves)
{
    const size_t waves = (n + particles_per_wave - 1) / particles_per_wave;
    size_t blocks = (waves + block_waves - 1) / block_waves;
    blocks = (blocks + kXcds - 1) / kXcds * kXcds;
    return dim3((unsigned)blocks);
}

int Fluid3D::enqueue_tile_passes(hipStream_t s, int order)
{
    const uint32_t n = (uint32_t)n_;
    const float4 mn = to_f4(box_min), mx = to_f4(box_max), gf = gravity_force_of(gravity);
    auto* pos = pos_.as<float4>();
    auto* vel = vel_.as<float4>();
    auto* force = force_.as<float4>();
    auto* spos = spos_.as<float4>();
    auto* svel = svel_.as<float4>();
    auto* index = index_.as<uint32_t>();
    auto* cs = cell_start_.as<uint32_t>();
    auto* ce = cell_end_.as<uint32_t>();
    auto* inv = inv_.as<uint32_t>();
    auto* porder = porder_.as<uint32_t>();
    const uint32_t lanes = (uint32_t)density_lanes;  // 2, 4 or 8 (set_tuning checks)
    const dim3 b(block_waves * kWave);
    const dim3 gd = tile_grid(n, kWave / lanes, block_waves), g = tile_grid(n, kWave / kGroupLanes, block_waves);
    const size_t lds_d = tile_lds_bytes(false, block_waves, tile_depth, lanes);
    const size_t lds_f = tile_lds_bytes(true, block_waves, tile_depth);
    const uint32_t depth = tile_depth;
    const float cut = sqrt_cut(params.cellSize);
    uint32_t* counts_out = record_counts ? counts_.as<uint32_t>() : nullptr;
    timer.begin(MFL_STAGE_DENSITY, s);
    auto* nbr = nbr_.as<uint32_t>();
    auto* ncount = ncount_.as<uint32_t>();
    auto* kept = kept_.as<uint32_t>();
    auto* hitmask = hitmask_.as<uint64_t>();
    auto* hitx = hitx_.as<uint64_t>();
    auto* meta = meta_.as<uint32_t>();
    const int out = neighbour_pass == MFL_PASS_TILE_LIST ? kOutList
                    : neighbour_pass == MFL_PASS_TILE_KEPT ? kOutKept : kOutNone;
#define MFL_DENSITY_TILE(O, OUT, W)                                                                              \
    hipLaunchKernelGGL((k_density_tile<O, OUT, W>), gd, b, lds_d, s, pos, spos, svel, cs, ce, n, params, depth,   \
                       porder, inv, nbr, ncount, kept, hitmask, hitx, meta, cut, counts_out)
#define MFL_DENSITY_TILE_W(O, OUT)                 \
    if (lanes == 2) MFL_DENSITY_TILE(O, OUT, 16); \
    else if (lanes == 4) MFL_DENSITY_TILE(O, OUT, 32); \
    else MFL_DENSITY_TILE(O, OUT, 64)
#define MFL_DENSITY_TILE_OUT(O)                                       \
    if (out == kOutList) { MFL_DENSITY_TILE_W(O, kOutList); }         \
    else if (out == kOutKept) { MFL_DENSITY_TILE_W(O, kOutKept); }    \
    else { MFL_DENSITY_TILE_W(O, kOutNone); }
    switch (order) {
    case MFL_TRAVERSAL_SORTED: MFL_DENSITY_TILE_OUT(kOrderSorted); break;
    case MFL_TRAVERSAL_CALLER: MFL_DENSITY_TILE_OUT(kOrderCaller); break;
    default: MFL_DENSITY_TILE_OUT(kOrderListed);
    }
#undef MFL_DENSITY_TILE_OUT
#undef MFL_DENSITY_TILE_W
#undef MFL_DENSITY_TILE
    timer.end(MFL_STAGE_DENSITY, s);
    if (out == kOutList) return enqueue_list_forces(s, order);
    timer.begin(MFL_STAGE_FORCES, s);
    if (out == kOutKept) {
        const size_t lds_k = (size_t)block_waves * (8 / lanes) * (lanes < 8 ? 32 : 64) * 2 * sizeof(float4);
#define MFL_FORCES_KEPT(O, W, STEP)                                                                              \
    hipLaunchKernelGGL((k_forces_kept<O, W, STEP>), gd, b, lds_k, s, spos, svel, pos, vel, force, index, cs, ce, n, \
                       params, depth, mn, mx, gf, porder, inv, kept, hitmask, hitx, meta)
#define MFL_FORCES_KEPT_W(O, STEP)                 \
    if (lanes == 2) MFL_FORCES_KEPT(O, 16, STEP); \
    else if (lanes == 4) MFL_FORCES_KEPT(O, 32, STEP); \
    else MFL_FORCES_KEPT(O, 64, STEP)
#define MFL_FORCES_KEPT_O(STEP)                                             \
    switch (order) {                                                        \
    case MFL_TRAVERSAL_SORTED: MFL_FORCES_KEPT_W(kOrderSorted, STEP); break; \
    case MFL_TRAVERSAL_CALLER: MFL_FORCES_KEPT_W(kOrderCaller, STEP); break; \
    default: MFL_FORCES_KEPT_W(kOrderListed, STEP);                         \
    }
        if (integrate_cb) {
            MFL_FORCES_KEPT_O(false);
        } else {
            MFL_FORCES_KEPT_O(true);
        }
#undef MFL_FORCES_KEPT_O
#undef MFL_FORCES_KEPT_W
#undef MFL_FORCES_KEPT
        timer.end(MFL_STAGE_FORCES, s);
        MFL_HIP(hipGetLastError());
        if (integrate_cb) (void)integrate_cb(handle, (void*)s, integrate_user);  // return value ignored, cpp:325-328
        return 1;
    }
#define MFL_FORCES_TILE(O, STEP)                                                                                  \
    hipLaunchKernelGGL((k_forces_tile<O, STEP>), g, b, lds_f, s, spos, svel, pos, vel, force, index, cs, ce, n,    \
                       params, depth, mn, mx, gf, porder, inv)
    if (integrate_cb) {
        switch (order) {
        case MFL_TRAVERSAL_SORTED: MFL_FORCES_TILE(kOrderSorted, false); break;
        case MFL_TRAVERSAL_CALLER: MFL_FORCES_TILE(kOrderCaller, false); break;
        default: MFL_FORCES_TILE(kOrderListed, false);
        }
        timer.end(MFL_STAGE_FORCES, s);
        MFL_HIP(hipGetLastError());
        (void)integrate_cb(handle, (void*)s, integrate_user);  // return value ignored, cpp:325-328
        return 1;
    }
    switch (order) {
    case MFL_TRAVERSAL_SORTED: MFL_FORCES_TILE(kOrderSorted, true); break;
    case MFL_TRAVERSAL_CALLER: MFL_FORCES_TILE(kOrderCaller, true); break;
    default: MFL_FORCES_TILE(kOrderListed, true);
    }
#undef MFL_FORCES_TILE
    timer.end(MFL_STAGE_FORCES, s);
    MFL_HIP(hipGetLastError());
    return 1;
}

int Fluid3D::get_tuning(mfl_tuning* out) const
{
    if (!out) {
        set_error("null tuning");
        return 0;
    }
    out->neighbour_pass = neighbour_pass;
    out->density_lanes = density_lanes;
    out->tile_depth = tile_depth;
    out->block_waves = block_waves;
    out->sort_items = (uint32_t)sort.tuning.items;
    out->force_chunks = force_chunks;
    return 1;
}

int Fluid3D::set_tuning(const mfl_tuning& t)
{
    if (t.neighbour_pass != MFL_PASS_LIST && t.neighbour_pass != MFL_PASS_TILE &&
        t.neighbour_pass != MFL_PASS_TILE_LIST && t.neighbour_pass != MFL_PASS_TILE_KEPT) {
        set_error("neighbour_pass must be MFL_PASS_LIST, MFL_PASS_TILE, MFL_PASS_TILE_LIST or MFL_PASS_TILE_KEPT");
        return 0;
    }
    if (t.neighbour_pass == MFL_PASS_LIST ? (t.density_lanes != 1 && t.density_lanes != 8)
                                          : (t.density_lanes != 2 && t.density_lanes != 4 && t.density_lanes != 8)) {
        set_error(t.neighbour_pass == MFL_PASS_LIST ? "density_lanes must be 1 or 8 with MFL_PASS_LIST"
                                                    : "density_lanes must be 2, 4 or 8 with the tile passes");
        return 0;
    }
    if (t.tile_depth < kTileMinDepth || t.tile_depth > kTileMaxDepth || t.tile_depth % 16 != 0) {
        set_error("tile_depth must be a multiple of 16 in [64, 512]");
        return 0;
    }
    if (t.block_waves != 1 && t.block_waves != 2 && t.block_waves != 4) {
        set_error("block_waves must be 1, 2 or 4");
        return 0;
    }
    if (t.neighbour_pass != MFL_PASS_LIST && t.tile_depth > 64u * (uint32_t)t.density_lanes) {
        set_error("tile_depth above 64 x density_lanes is not available with the tile passes");
        return 0;
    }
    if (tile_lds_bytes(true, t.block_waves, t.tile_depth) > 65536 ||
        (t.neighbour_pass != MFL_PASS_LIST &&
         tile_lds_bytes(false, t.block_waves, t.tile_depth, (uint32_t)t.density_lanes) > 65536)) {
        set_error("tile_depth x block_waves needs more than 64 KiB of LDS per workgroup");
        return 0;
    }
    if (t.sort_items != 0 && t.sort_items != 4 && t.sort_items != 8 && t.sort_items != 16) {
        set_error("sort_items must be 0, 4, 8 or 16");
        return 0;
    }
    if (t.force_chunks != 1 && t.force_chunks != 2) {
        set_error("force_chunks must be 1 or 2");
        return 0;
    }
    neighbour_pass = t.neighbour_pass;
    density_lanes = t.density_lanes;
    tile_depth = t.tile_depth;
    block_waves = t.block_waves;
    sort.tuning.items = (int)t.sort_items;
    force_chunks = t.force_chunks;
    return 1;
}

int Fluid3D::autotune(hipStream_t s, uint32_t steps, mfl_tuning* best, double* ms_per_step)
{
    if (!check_user_buffers()) return 0;
    if (steps == 0) {
        set_error("autotune needs at least one step per trial");
        return 0;
    }
    // the trials re-size the sort's temporaries (other tile sizes): no graph may keep the old ones
    drop_graph();
    // snapshot of everything a step reads that a step also writes
    const size_t vb = n_ * sizeof(float4), tb = (size_t)kBuckets * sizeof(uint32_t);
    DeviceBuffer save;
    if (!save.allocate(3 * vb + 2 * tb)) return 0;
    char* sp = save.as<char>();
    auto snapshot = [&](bool restore) -> int {
        void* dev[] = {pos_.ptr, vel_.ptr, force_.ptr, cell_start_.ptr, cell_end_.ptr};
        const size_t len[] = {vb, vb, vb, tb, tb};
        size_t off = 0;
        for (int k = 0; k < 5; k++) {
            void* dst = restore ? dev[k] : sp + off;
            const void* src = restore ? (const void*)(sp + off) : dev[k];
            MFL_HIP(hipMemcpyAsync(dst, src, len[k], hipMemcpyDeviceToDevice, s));
            off += len[k];
        }
        return 1;
    };
    if (!snapshot(false)) {
        save.release();
        return 0;
    }
    mfl_tuning start{};
    get_tuning(&start);
    std::vector<mfl_tuning> cand;
    for (uint32_t ch : {1u, 2u})
        for (int lanes : {1, 8}) cand.push_back({MFL_PASS_LIST, lanes, 256u, 4u, start.sort_items, ch});
    for (uint32_t w : {1u, 2u, 4u})
        for (uint32_t d : {96u, 128u, 192u, 256u, 384u, 512u})
            for (int lanes : {2, 4, 8})
                if (tile_lds_bytes(true, w, d) <= 65536 && d <= 64u * (uint32_t)lanes) {
                    cand.push_back({MFL_PASS_TILE, lanes, d, w, start.sort_items, 1u});
                    cand.push_back({MFL_PASS_TILE_KEPT, lanes, d, w, start.sort_items, 1u});
                    for (uint32_t ch : {1u, 2u}) cand.push_back({MFL_PASS_TILE_LIST, lanes, d, w, start.sort_items, ch});
                }
    const bool graph_was = use_graph;
    const uint32_t mask_was = timer.mask;
    const uint32_t since = steps_since_refresh_;
    use_graph = false;
    timer.mask = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int ok = hip_ok(hipEventCreate(&e0), "hipEventCreate") && hip_ok(hipEventCreate(&e1), "hipEventCreate");
    // ms per step of shape t over k timed steps from the snapshot (after one untimed step), or a negative
    // value when a call failed
    auto trial = [&](const mfl_tuning& t, uint32_t k) -> double {
        ok = set_tuning(t) && snapshot(true) && compute(s);
        for (uint32_t j = 0; ok && j < k; j++) {
            if (j == 0) ok = hip_ok(hipEventRecord(e0, s), "hipEventRecord");
            ok = ok && compute(s);
        }
        ok = ok && hip_ok(hipEventRecord(e1, s), "hipEventRecord") && hip_ok(hipEventSynchronize(e1), "sync");
        float ms = 0.f;
        ok = ok && hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ok ? (double)ms / k : -1.0;
    };
    // The card lowers its clocks within milliseconds of idling and needs tens of milliseconds of work to
    // raise them: step the starting shape untimed for 30 ms first, so that the first candidates are not
    // timed at lower clocks than the last ones.
    {
        const auto t0 = std::chrono::steady_clock::now();
        ok = ok && set_tuning(start);
        while (ok && std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(30)) {
            ok = snapshot(true);
            for (int j = 0; ok && j < 16; j++) ok = compute(s);
            ok = ok && hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
        }
    }
    // every candidate for `steps` steps, then the four fastest again for four times as many (a short
    // trial's ranking is noisy), then the sort's tile size with the best neighbour passes
    std::vector<std::pair<double, size_t>> times;
    for (size_t c = 0; ok && c < cand.size(); c++) {
        const double ms = trial(cand[c], steps);
        if (ok) times.emplace_back(ms, c);
    }
    std::sort(times.begin(), times.end());
    double best_ms = 1e30;
    mfl_tuning pick = start;
    for (size_t r = 0; ok && r < times.size() && r < 4; r++) {
        const double ms = trial(cand[times[r].second], 4 * steps);
        if (ok && ms < best_ms) {
            best_ms = ms;
            pick = cand[times[r].second];
        }
    }
    const mfl_tuning passes = pick;
    for (uint32_t items : {4u, 8u, 16u}) {
        mfl_tuning t = passes;
        t.sort_items = items;
        if (!ok || items == passes.sort_items) continue;
        const double ms = trial(t, 4 * steps);
        if (ok && ms < best_ms) {
            best_ms = ms;
            pick = t;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    use_graph = graph_was;
    timer.mask = mask_was;
    const int restored = snapshot(true) && hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
    steps_since_refresh_ = since;
    porder_valid_ = false;  // the trials moved the particles; rebuild the traversal order on the next step
    set_tuning(ok ? pick : start);
    save.release();
    if (!ok || !restored) return 0;
    if (best) *best = pick;
    if (ms_per_step) *ms_per_step = best_ms;
    return 1;
}

void Fluid3D::drop_graph()
{
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_sig_.clear();
}

int Fluid3D::launch_graph(hipStream_t s)
{
    // the sort's temporaries are reserved first: the signature holds their addresses
    if (!sort.reserve(n_, 18)) return 0;
    std::string sig;
    sig.reserve(256);
    sort.sig_append(sig);
    sig_add(sig, n_);
    sig_add(sig, pos_.ptr);
    sig_add(sig, vel_.ptr);
    sig_add(sig, force_.ptr);
    sig_add(sig, clear_count);
    sig_add(sig, sort.tuning.items);
    sig_add(sig, traversal);
    sig_add(sig, porder_valid_);
    sig_add(sig, density_lanes);
    sig_add(sig, neighbour_pass);
    sig_add(sig, tile_depth);
    sig_add(sig, block_waves);
    sig_add(sig, force_chunks);
    sig_add(sig, record_counts);
    sig_add(sig, params);
    sig_add(sig, box_min);
    sig_add(sig, box_max);
    sig_add(sig, gravity);
    if (!graph_exec_ || sig != graph_sig_) {
        drop_graph();
        MFL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        const int ok = enqueue_step(s);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (!ok || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (ok) hip_ok(e, "hipStreamEndCapture");
            return 0;
        }
        graph_ = g;
        MFL_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        graph_sig_ = sig;
    }
    MFL_HIP(hipGraphLaunch(graph_exec_, s));
    return 1;
}

// Bits per axis of the traversal's Morton keys: 8 when the box lies in [0, 254 h] on every axis (its
// cells' coordinates are below 256, where 8-bit keys order the cells exactly as 10-bit ones: 24-bit
// keys, three sort passes instead of four), else 10. Any key is a valid order; a particle outside the
// box only gets a less local one.
uint32_t Fluid3D::morton_bits() const
{
    const float h = params.cellSize;
    if (!(h > 0.f)) return 10;
    for (int a = 0; a < 3; a++)
        if (!(box_min[a] >= 0.f) || !(box_max[a] <= 254.f * h)) return 10;
    return 8;
}

int Fluid3D::refresh_order(hipStream_t s)
{
    const uint32_t n = (uint32_t)n_;
    const uint32_t mbits = morton_bits();
    timer.begin(MFL_STAGE_ORDER, s);
    hipLaunchKernelGGL(k_morton_keys, grid_for(n), dim3(kThreads), 0, s, pos_.as<float4>(), n, params.cellSize, mbits,
                       mkey_.as<uint32_t>());
    MFL_HIP(hipGetLastError());
    if (!sort.sort_bits(mkey_.as<uint32_t>(), porder_.as<uint32_t>(), n, 0, (int)(3 * mbits), s, true)) return 0;
    timer.end(MFL_STAGE_ORDER, s);
    porder_valid_ = true;
    return 1;
}

int Fluid3D::compute(hipStream_t s)
{
    if (!check_user_buffers()) return 0;
    if (traversal == MFL_TRAVERSAL_MORTON &&
        (!porder_valid_ || (refresh_interval > 0 && steps_since_refresh_ >= refresh_interval))) {
        if (!refresh_order(s)) return 0;
        steps_since_refresh_ = 0;
    }
    steps_since_refresh_++;
    if (use_graph && !integrate_cb && s != nullptr && timer.mask == 0) return launch_graph(s);
    return enqueue_step(s);
}

int Fluid3D::enqueue_kernel(const char* name, hipStream_t s)
{
    if (!name) {
        set_error("null kernel name");
        return 0;
    }
    if (!check_user_buffers()) return 0;
    const uint32_t n = (uint32_t)n_;
    const dim3 g = grid_for(n), b(kThreads);
    const float4 mn = to_f4(box_min), mx = to_f4(box_max), gf = gravity_force_of(gravity);
    auto* pos = pos_.as<float4>();
    auto* vel = vel_.as<float4>();
    auto* force = force_.as<float4>();
    auto* spos = spos_.as<float4>();
    auto* svel = svel_.as<float4>();
    auto* cell = cell_.as<uint32_t>();
    auto* index = index_.as<uint32_t>();
    auto* cs = cell_start_.as<uint32_t>();
    auto* ce = cell_end_.as<uint32_t>();
    const std::string k(name);
    if (k == "clHash")
        hipLaunchKernelGGL((k_hash<true, false>), g, b, 0, s, pos, n, params.cellSize, cell, index,
                           params_buf_.as<mfl_params>(), params, nullptr, nullptr, 0u);
    else if (k == "clReorder")
        hipLaunchKernelGGL(k_reorder_clear, g, b, 0, s, index, pos, vel, spos, svel, n, cs, ce, 0u, inv_.as<uint32_t>());
    else if (k == "clInitBounds")
        hipLaunchKernelGGL(k_fill_buckets, grid_for(kBuckets), b, 0, s, cs, ce, kBuckets);
    else if (k == "clFindBounds")
        hipLaunchKernelGGL(k_find_bounds, g, b, 0, s, cell, n, cs, ce);
    else if (k == "clCalculateDensity")
        hipLaunchKernelGGL(k_density, g, b, 0, s, spos, svel, cs, ce, n, params,
                           record_counts ? counts_.as<uint32_t>() : nullptr);
    else if (k == "clCalculateForces")
        hipLaunchKernelGGL(k_forces, g, b, 0, s, spos, svel, force, index, cs, ce, n, params);
    else if (k == "clClipBox")
        hipLaunchKernelGGL(k_clip_box, g, b, 0, s, spos, svel, force, index, n, params, mn, mx);
    else if (k == "clIntegrateForce")
        hipLaunchKernelGGL(k_integrate_force, g, b, 0, s, vel, force, n, params);
    else if (k == "clGravity")
        hipLaunchKernelGGL(k_gravity, g, b, 0, s, force, n, gf);
    else if (k == "clIntegrateVelocity")
        hipLaunchKernelGGL(k_integrate_velocity, g, b, 0, s, pos, vel, n, params);
    else {
        set_error("no kernel named " + k);  // KERNEL_VALIDATE, libCL/oclKernel.h:89-94
        return 0;
    }
    MFL_HIP(hipGetLastError());
    return 1;
}

}  // namespace mfl
