// Batched scenes on gfx950: S independent oclFluid3D systems (libCL/phys/oclFluid3D.cpp:293-342 each)
// advanced by one set of launches per step, optionally replayed from one hipGraph.
//
// Layout: the scenes' particles are concatenated scene by scene (offsets off[s], counts multiples of
// 256); each scene keeps its own pair of 2^18-entry bucket tables, holding scene-local indices exactly
// as a single-scene program would. Keys carry the scene in their high bits, key = s << 18 | hash, so
// one stable sort of the concatenation is the segmented sort: it orders every scene by hash and keeps
// the scenes in place. Every per-particle computation reuses the single-scene arithmetic
// (sph_device.hpp) on scene-local pointers, so each scene's state is bit-identical to running it alone.
#include "batch.hpp"
#include "sph_device.hpp"
#include "tile.hpp"

#include <chrono>
#include <cstring>

namespace mfl {
namespace {

using namespace sph;

// Keys s << 18 | hash per particle, and the sort's first-pass digit counts (RadixSort::first_pass_plan):
// block b hashes the keys of sort tile xcd_tile(), ITEMS per thread at tile offsets k * 256 + thread,
// and histograms their low `bits` bits in LDS. The grid is one block per tile. The same launch resets
// the first `clear` entries of every scene's tables (cpp:309-313; the reference resets them between the
// reorder and the bounds, and nothing before the bounds kernel reads them): block b takes its share of
// the S x clear entries, 16 bytes per store when clear is a multiple of 4.
template <int ITEMS>
__global__ __launch_bounds__(kThreads) void k_b_hash_counts(const float4* __restrict__ pos,
                                                            const uint32_t* __restrict__ pscene,
                                                            const mfl_params* __restrict__ params, uint32_t N,
                                                            uint32_t* __restrict__ cell, uint32_t* __restrict__ counts,
                                                            uint32_t ntiles, uint32_t bits,
                                                            uint32_t* __restrict__ cs_all,
                                                            uint32_t* __restrict__ ce_all, uint32_t clear, uint32_t S)
{
    __shared__ uint32_t hist[512];
    const uint32_t radix = 1u << bits, mask = radix - 1u;
    for (uint32_t d = threadIdx.x; d < radix; d += kThreads) hist[d] = 0;
    __syncthreads();
    const uint32_t tile = xcd_tile();
    const size_t base = (size_t)tile * (kThreads * ITEMS);
    // every load issued before the first count (a clamped index keeps the loads unconditional)
    float4 q[ITEMS];
    uint32_t sc[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const size_t g = base + (size_t)k * kThreads + threadIdx.x;
        const size_t gc = g < N ? g : N - 1;
        q[k] = pos[gc];
        sc[k] = pscene[gc];
    }
    if (clear) {  // while the loads are in flight
        const uint32_t unit = (clear & 3u) == 0u ? 4u : 1u, per = clear / unit;  // units per scene
        const size_t total = (size_t)S * per, chunk = (total + gridDim.x - 1) / gridDim.x;
        const size_t lo = (size_t)blockIdx.x * chunk, hi = lo + chunk < total ? lo + chunk : total;
        size_t u = lo + threadIdx.x;
        uint32_t s = (uint32_t)(u / per), k = (uint32_t)(u % per);
        for (; u < hi; u += kThreads) {
            const size_t at = (size_t)s * kBuckets + (size_t)k * unit;
            if (unit == 4u) {
                const uint4 e = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
                *reinterpret_cast<uint4*>(cs_all + at) = e;
                *reinterpret_cast<uint4*>(ce_all + at) = e;
            } else {
                cs_all[at] = kEmpty;
                ce_all[at] = kEmpty;
            }
            for (k += kThreads; k >= per; k -= per) s++;
        }
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const size_t g = base + (size_t)k * kThreads + threadIdx.x;
        if (g < N) {
            const uint32_t s = sc[k];
            const float h = params[s].cellSize;
            const uint32_t key =
                (s << 18) | cell_hash(grid_coord(q[k].x, h), grid_coord(q[k].y, h), grid_coord(q[k].z, h));
            cell[g] = key;
            atomicAdd(&hist[key & mask], 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < radix; d += kThreads) counts[(size_t)d * ntiles + tile] = hist[d];
}

// The sorted state gathered through the index (clReorder, cl:66-72), the inverse permutation
// (inv[index[i]] = i, for the neighbour passes' traversal) and clFindBounds (cl:85-113) per scene in
// scene-local indices (the tables were reset before the sort; nothing between reads them). The gather
// reads inside one scene's range (at most 65,536 particles, 2 MiB of pos + vel): blocks are dealt to
// the XCDs in contiguous ranges of sorted slots (xcd_tile), so the blocks of one scene share an XCD's L2.
__global__ __launch_bounds__(kThreads) void k_b_slots_bounds(const uint32_t* __restrict__ index, uint32_t N,
                                                             uint32_t* __restrict__ inv,
                                                             const uint32_t* __restrict__ cell,
                                                             const uint32_t* __restrict__ off,
                                                             uint32_t* __restrict__ cs_all,
                                                             uint32_t* __restrict__ ce_all,
                                                             const float4* __restrict__ pos,
                                                             const float4* __restrict__ vel,
                                                             float4* __restrict__ spos, float4* __restrict__ svel)
{
    const uint32_t i = xcd_tile() * kThreads + threadIdx.x;
    if (i >= N) return;
    const uint32_t g = index[i];
    if (g < N) {  // index is a permutation of [0, N)
        inv[g] = i;
        spos[i] = pos[g];
        svel[i] = vel[g];
    }
    const uint32_t key = cell[i];
    const uint32_t s = key >> 18, h = key & (kBuckets - 1u);
    const uint32_t lo = off[s], n_s = off[s + 1] - lo, li = i - lo;
    uint32_t* cs = cs_all + (size_t)s * kBuckets;
    uint32_t* ce = ce_all + (size_t)s * kBuckets;
    if (li == 0) {
        cs[h] = 0;
        return;
    }
    const uint32_t prev = cell[i - 1] & (kBuckets - 1u);
    if (h != prev) {
        ce[prev] = li;
        cs[h] = li;
    }
    if (li == n_s - 1) ce[h] = n_s;
}

// Traversal order of the tile density pass: scenes in order (scene in the high key bits), a Morton
// order of the cells inside each scene (mbits bits per axis, wrapped).
__global__ __launch_bounds__(kThreads) void k_b_morton_keys(const float4* __restrict__ pos,
                                                            const uint32_t* __restrict__ pscene,
                                                            const mfl_params* __restrict__ params, uint32_t N,
                                                            uint32_t mbits, uint32_t* __restrict__ key)
{
    const uint32_t g = blockIdx.x * kThreads + threadIdx.x;
    if (g >= N) return;
    const uint32_t s = pscene[g];
    const float h = params[s].cellSize;
    const float4 q = pos[g];
    const uint32_t m = (1u << mbits) - 1u;
    const uint32_t code = spread_bits3((uint32_t)grid_coord(q.x, h) & m) |
                          (spread_bits3((uint32_t)grid_coord(q.y, h) & m) << 1) |
                          (spread_bits3((uint32_t)grid_coord(q.z, h) & m) << 2);
    key[g] = (s << (3 * mbits)) | code;
}

// One group of eight lanes per particle (k_density_grp of the single-scene program) on scene s.
__device__ inline void batch_density_home(uint32_t s, int gx, int gy, int gz, const float4& pi, uint32_t i, uint32_t slot,
                                          uint32_t grp_shift, bool live, const float4* __restrict__ spos,
                                          const uint32_t* __restrict__ cs_all, const uint32_t* __restrict__ ce_all,
                                          const uint32_t* __restrict__ off, const mfl_params* __restrict__ params,
                                          uint32_t N, uint32_t* __restrict__ nbr, float& acc, uint32_t& count, float& h)
{
    const uint32_t lo = off[s], n_s = off[s + 1] - lo;
    h = params[s].cellSize;
    const float h2 = h * h;
    const float h2_lo = h2 * 0.9999996f, h2_hi = h2 * 1.0000005f;
    density_visits(gx, gy, gz, pi, i - lo, i, slot, grp_shift, live, spos + lo, cs_all + (size_t)s * kBuckets,
                   ce_all + (size_t)s * kBuckets, n_s, h, h2, h2_lo, h2_hi, nbr, N, acc, count);
}

__global__ __launch_bounds__(kThreads) void k_b_density(float4* __restrict__ spos, float4* __restrict__ svel,
                                                        const uint32_t* __restrict__ cs_all,
                                                        const uint32_t* __restrict__ ce_all,
                                                        const uint32_t* __restrict__ cell,
                                                        const uint32_t* __restrict__ off,
                                                        const mfl_params* __restrict__ params, uint32_t N,
                                                        uint32_t* __restrict__ nbr, uint32_t* __restrict__ ncount)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t grp = lane / kGroupLanes, slot = lane % kGroupLanes;
    const uint32_t wave = (blockIdx.x * kThreads + threadIdx.x) / kWave;
    const uint32_t i = wave * (kWave / kGroupLanes) + grp;  // sorted (scene-major) index
    const bool live = i < N;
    const uint32_t s = live ? cell[i] >> 18 : 0u;
    const float h0 = params[s].cellSize;
    const float4 pi = live ? spos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int gx = grid_coord(pi.x, h0), gy = grid_coord(pi.y, h0), gz = grid_coord(pi.z, h0);
    const uint64_t live_mask = __ballot(live);
    if (live_mask == 0) return;
    const int src = __ffsll((unsigned long long)live_mask) - 1;
    const uint32_t s0 = __shfl(s, src, kWave);
    const int gx0 = __shfl(gx, src, kWave), gy0 = __shfl(gy, src, kWave), gz0 = __shfl(gz, src, kWave);
    const bool same = !live || (s == s0 && gx == gx0 && gy == gy0 && gz == gz0);
    float acc = 0.f, h = h0;
    uint32_t count = 0;
    const uint32_t grp_shift = grp * kGroupLanes;
    if (__ballot(!same) == 0)
        batch_density_home(__builtin_amdgcn_readfirstlane(s0), __builtin_amdgcn_readfirstlane(gx0),
                           __builtin_amdgcn_readfirstlane(gy0), __builtin_amdgcn_readfirstlane(gz0), pi, i, slot,
                           grp_shift, live, spos, cs_all, ce_all, off, params, N, nbr, acc, count, h);
    else
        batch_density_home(s, gx, gy, gz, pi, i, slot, grp_shift, live, spos, cs_all, ce_all, off, params, N, nbr, acc,
                           count, h);
    acc += __shfl_xor(acc, 1, kWave);
    acc += __shfl_xor(acc, 2, kWave);
    acc += __shfl_xor(acc, 4, kWave);
    if (!live || slot != 0) return;
    const mfl_params& p = params[s];
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
    ncount[i] = count;
}

// One lane walks particle i's 27 bucket ranges of its scene from global memory and records the visits
// (the overflow path of the batched tile pass; scene-local indices throughout).
// Sums returned by value: a pointer to the caller's accumulators would keep them in scratch memory.
struct BVisitSums {
    float acc;
    uint32_t count;
};

__device__ __noinline__ BVisitSums b_density_walk(int cx, int cy, int cz, float4 pi, uint32_t il, uint32_t ig,
                                                 const float4* __restrict__ spos, const uint32_t* __restrict__ cs,
                                                 const uint32_t* __restrict__ ce, uint32_t n_s, float h, uint32_t N,
                                                 uint32_t* __restrict__ nbr)
{
    const float h2 = h * h;
    const float h2_lo = h2 * 0.9999996f, h2_hi = h2 * 1.0000005f;
    float acc = 0.f;
    uint32_t count = 0;
    for (int k = 0; k < 27; k++) {
        const uint32_t b = cell_hash(cx + k % 3 - 1, cy + (k / 3) % 3 - 1, cz + k / 9 - 1);
        const uint32_t st = cs[b];
        if (st == kEmpty) continue;
        const uint32_t end = min(ce[b], n_s);
        for (uint32_t j = st; j < end; j++) {
            const float4 pj = spos[j];
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float r2 = dist2(dx, dy, dz);
            if (j != il && accept_r2(r2, h, h2_lo, h2_hi)) {
                const float w = h2 - r2;
                acc = fmaf(w * w, w, acc);
                if (nbr && count < kNbrCap) nbr[nbr_at(ig, count, N)] = j;
                count++;
            }
        }
    }
    return {acc, count};
}

// The tile density pass of the single-scene program (k_density_tile) over all scenes: traversal slot t
// in the batch's Morton order, each 8-particle segment inside one scene (scene counts are multiples of
// 256 and the order keeps scenes in place), that scene's Params, tables and scene-local indices.
// KEPT = false records the visits by sorted slot for k_b_forces_step (TILE_LIST); KEPT = true writes
// each cell's kept entries, the lanes' acceptance bits and the per-particle meta for k_b_forces_kept
// (TILE_KEPT), with the same layout as the single-scene pass.
template <int W, bool KEPT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_b_density_tile(const float4* __restrict__ pos, float4* __restrict__ spos,
                                                             float4* __restrict__ svel,
                                                             const uint32_t* __restrict__ cs_all,
                                                             const uint32_t* __restrict__ ce_all,
                                                             const uint32_t* __restrict__ off,
                                                             const uint32_t* __restrict__ pscene,
                                                             const mfl_params* __restrict__ params, uint32_t N,
                                                             uint32_t depth, const uint32_t* __restrict__ porder,
                                                             const uint32_t* __restrict__ inv,
                                                             uint32_t* __restrict__ nbr, uint32_t* __restrict__ ncount,
                                                             uint32_t* __restrict__ kept_out,
                                                             uint64_t* __restrict__ hitmask, uint64_t* __restrict__ hitx,
                                                             uint32_t* __restrict__ meta, const float* __restrict__ cuts)
{
    constexpr uint32_t P = W / 8;
    constexpr uint32_t SEGS = kWave / W;
    extern __shared__ float4 lds[];
    const Lanes<W> L(threadIdx.x & (kWave - 1));
    const uint32_t nseg = (blockDim.x / kWave) * SEGS, sid = (threadIdx.x / kWave) * SEGS + L.seg;
    float4* tile = lds + (size_t)sid * (depth + kGroupLanes);
    StageAux* aux = reinterpret_cast<StageAux*>(lds + (size_t)nseg * (depth + kGroupLanes)) + sid;
    const uint32_t grp = L.ls / P, slot = L.ls % P;
    const uint32_t wave = (xcd_block() * blockDim.x + threadIdx.x) / kWave;
    const uint32_t t = wave * (kWave / P) + L.seg * 8 + grp;
    uint32_t i = N, sc = 0;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (t < N) {
        const uint32_t o = porder[t];
        if (o < N) {
            i = inv[o];
            pi = pos[o];
            sc = pscene[o];
        }
    }
    const bool live = i < N;
    // The wave's scene, in scalar registers: a wave holds 16 (or 8) consecutive traversal slots from a
    // multiple of 16, scenes hold multiples of 256 slots in the traversal order, so a wave never
    // straddles two scenes (and is all live or all past N, whose lanes carry scene 0).
    const uint32_t s = __builtin_amdgcn_readfirstlane(sc);
    const uint32_t lo = off[s], n_s = off[s + 1] - lo;
    const float4* sp = spos + lo;
    const uint32_t* cs = cs_all + (size_t)s * kBuckets;
    const uint32_t* ce = ce_all + (size_t)s * kBuckets;
    const uint32_t il = live ? i - lo : n_s;
    const float h = params[s].cellSize;  // the other Params are read after the loop (fewer live VGPRs)
    const float h2 = h * h;
    const float cut = cuts[s];  // sqrt(r2) < h exactly when r2 < cut
    const int gx = grid_coord(pi.x, h), gy = grid_coord(pi.y, h), gz = grid_coord(pi.z, h);
    float acc = 0.f;
    uint32_t count = 0;
    uint32_t used = 0;  // KEPT: entries of this segment's kept buffer already written
    const uint32_t gseg = wave * SEGS + L.seg;
    uint64_t todo = L.bits(__ballot(live));
    while (__ballot(todo != 0) != 0) {
        const SegCell c = next_cell_seg<W>(L, todo, live, gx, gy, gz, pi);
        todo &= ~L.bits(__ballot(c.member));
        const uint32_t total = stage_seg_ranges<W>(L, c, cs, ce, n_s, aux);
        const uint32_t kept = total == kStageOverflow ? kStageOverflow
                                                      : stage_seg_fill<W>(L, total, c.lo, c.hi, h2 * kKeepMargin, sp, 0u, depth, tile, aux);
        if constexpr (KEPT) {
            // the cell's kept entries go after the segment's earlier cells when they all fit
            if (kept != kStageOverflow && used + kept > kSegCap && c.member && slot == 0) meta[t] = kMetaWalk;
        }
        uint32_t at_base = 0;  // !KEPT: the particle's visits recorded in earlier windows
        const bool fits = sweep_windows<W>(
            L, kept, h2 * kKeepMargin, sp, depth, tile, aux, pi, il, h2, cut,
            [&](uint32_t w, uint32_t nw, uint32_t trips, uint64_t hits, float a) {
                if (c.member) {
                    acc += a;
                    count += (uint32_t)__popcll(hits);
                }
                if constexpr (KEPT) {
                    if (used + kept <= kSegCap) {
                        for (uint32_t q = L.ls; q < nw; q += W)
                            kept_out[(size_t)gseg * kSegCap + used + w * depth + q] = __float_as_uint(tile[q].w);
                        if (c.member) {
                            if (w == 0) hitmask[(size_t)t * P + slot] = hits;
                            else hitx[hitx_at(t, P, slot, w)] = hits;
                        }
                    }
                } else if (c.member) {
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
                        if (at < kNbrCap) nbr[nbr_at(i, at, N)] = __float_as_uint(tile[m * P + slot].w);
                        at++;
                    }
                }
            });
        if constexpr (KEPT) {
            if (fits && used + kept <= kSegCap) {
                if (c.member && slot == 0) meta[t] = used | (kept << 16);
                used += kept;
            }
        }
        if (!fits && c.member) {
            if constexpr (KEPT) {  // the particle's P lanes split the walk (density_walk_split, as alone)
                if (slot == 0) meta[t] = kMetaWalk;
                const WalkSums f = density_walk_split(c.cx, c.cy, c.cz, pi, il, sp, cs, ce, n_s, h, slot, P);
                count += f.count;
                acc += f.acc;
            } else if (slot == 0) {
                const BVisitSums f = b_density_walk(c.cx, c.cy, c.cz, pi, il, i, sp, cs, ce, n_s, h, N, nbr);
                count += f.count;
                acc += f.acc;
            }
        }
        wave_sync();
    }
    acc = particle_sum<W>(acc);
    count = particle_sum_u<W>(count);
    if (!live || slot != 0) return;
    const mfl_params p = params[s];
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
    if (!KEPT) ncount[i] = count;
}


// The kept-entry force pass of the single-scene program (k_forces_kept) over all scenes, with the same
// waves and segments as k_b_density_tile<W, true>: each segment stages its cell's kept entries (scene-
// local indices) 64 at a time in LDS and every lane sums the pair terms its acceptance bits mark,
// earliest trip first; then the scene's clip box, clamp, gravity and Euler update.
template <int W>
__global__ __launch_bounds__(kThreads) void k_b_forces_kept(const float4* __restrict__ spos,
                                                            const float4* __restrict__ svel, float4* __restrict__ pos,
                                                            float4* __restrict__ vel, float4* __restrict__ force,
                                                            const uint32_t* __restrict__ cs_all,
                                                            const uint32_t* __restrict__ ce_all,
                                                            const uint32_t* __restrict__ off,
                                                            const uint32_t* __restrict__ pscene,
                                                            const mfl_params* __restrict__ params,
                                                            const float4* __restrict__ boxes, uint32_t N,
                                                            uint32_t depth, const uint32_t* __restrict__ porder,
                                                            const uint32_t* __restrict__ inv,
                                                            const uint32_t* __restrict__ kept_in,
                                                            const uint64_t* __restrict__ hitmask,
                                                            const uint64_t* __restrict__ hitx,
                                                            const uint32_t* __restrict__ meta, float4 gravity_force)
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
    uint32_t i = N, o = N, sc = 0;
    if (t < N) {
        o = porder[t];
        if (o < N) {
            i = inv[o];
            sc = pscene[o];
        }
    }
    const bool live = i < N;
    const uint32_t s = __builtin_amdgcn_readfirstlane(sc);  // the wave's scene (k_b_density_tile)
    const uint32_t lo = off[s], n_s = off[s + 1] - lo;
    const float4* sp = spos + lo;
    const float4* sv = svel + lo;
    const float h = params[s].cellSize;  // the other Params are read after the loop (fewer live VGPRs)
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
        const uint32_t offk = active ? (gm & 0xFFFFu) : 0u, kept = active ? (gm >> 16) : 0u;
        const uint32_t kmax = seg_wave_max<W>(kept);
        // the density pass's windows of `depth` entries, each with its own word of acceptance bits
        for (uint32_t w0 = 0, w = 0; w0 < kmax; w0 += depth, w++) {
            const uint32_t nw = kept > w0 ? min(kept - w0, depth) : 0u, nmax = min(kmax - w0, depth);
            const uint32_t trips = (nw + P - 1) / P;
            const uint64_t hw = w == 0 ? hm : (member && nw ? hitx[hitx_at(t, P, slot, w)] : 0ull);
            const size_t kbase = (size_t)gseg * kSegCap + offk + w0;
            // the next chunk's index loads are in flight during a chunk's sweep
            uint32_t kidx[SEGS];
            kept_chunk_index<W>(kept_in, kbase, nw, 0u, L.lane, kidx);
            for (uint32_t c0 = 0; c0 < nmax; c0 += chunk) {
                kept_chunk_load<W>(kidx, nw, c0, L.lane, sp, sv, wpos, wvel);
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
                wave_sync();
            }
        }
    }
    if (live && !staged) {  // the particle's P lanes split the walk of a cell that overflowed its tile
        const WalkForces f = force_walk_split(sp, sv, cs_all + (size_t)s * kBuckets, ce_all + (size_t)s * kBuckets,
                                              i - lo, n_s, pi, vi, h, slot, P);
        Pa = f.P;
        Va = f.V;
    }
    Pa = make_float3(particle_sum<W>(Pa.x), particle_sum<W>(Pa.y), particle_sum<W>(Pa.z));
    Va = make_float3(particle_sum<W>(Va.x), particle_sum<W>(Va.y), particle_sum<W>(Va.z));
    if (!live || slot != 0) return;
    const mfl_params p = params[s];
    float4 f = add_sph_force(force[o], Pa, Va, p);
    f = clip_box(f, pi, vi, p.particleRadius, boxes[2 * s], boxes[2 * s + 1]);
    const float4 v = integrate_force(vel[o], f, p);
    force[o] = gravity_force;
    vel[o] = v;
    pos[o] = f4_add(pos[o], f4_scale(v, p.deltaTime));
}

template <int CH>
__global__ __launch_bounds__(kThreads) void k_b_forces_step(const float4* __restrict__ spos,
                                                            const float4* __restrict__ svel, float4* __restrict__ pos,
                                                            float4* __restrict__ vel, float4* __restrict__ force,
                                                            const uint32_t* __restrict__ index,
                                                            const uint32_t* __restrict__ cell,
                                                            const uint32_t* __restrict__ cs_all,
                                                            const uint32_t* __restrict__ ce_all,
                                                            const uint32_t* __restrict__ off,
                                                            const mfl_params* __restrict__ params,
                                                            const float4* __restrict__ boxes,
                                                            const uint32_t* __restrict__ nbr,
                                                            const uint32_t* __restrict__ ncount, uint32_t N,
                                                            float4 gravity_force)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    const uint32_t s = cell[i] >> 18;
    const uint32_t lo = off[s], n_s = off[s + 1] - lo;
    const mfl_params p = params[s];
    const float4 pi = spos[i], vi = svel[i];
    float3 P, V;
    const uint32_t count = ncount[i];
    if (count <= kNbrCap)
        sph_force_sum_nbr<CH>(spos + lo, svel + lo, nbr, N, count, i, pi, vi, p.cellSize, P, V);
    else
        sph_force_sum(spos + lo, svel + lo, cs_all + (size_t)s * kBuckets, ce_all + (size_t)s * kBuckets, i - lo, n_s,
                      pi, vi, p.cellSize, P, V);
    const uint32_t o = index[i];
    if (o >= N) return;
    float4 f = add_sph_force(force[o], P, V, p);
    f = clip_box(f, pi, vi, p.particleRadius, boxes[2 * s], boxes[2 * s + 1]);
    const float4 v = integrate_force(vel[o], f, p);
    force[o] = gravity_force;
    vel[o] = v;
    pos[o] = f4_add(pos[o], f4_scale(v, p.deltaTime));
}

}  // namespace

FluidBatch::FluidBatch(uint32_t num_scenes, const uint32_t* counts)
{
    if (num_scenes == 0 || num_scenes > (1u << 14) || !counts) {
        set_error("a batch holds between 1 and 16384 scenes");
        return;
    }
    S_ = num_scenes;
    off_host_.assign(S_ + 1, 0);
    uint64_t total = 0;
    for (uint32_t s = 0; s < S_; s++) {
        if (counts[s] == 0 || counts[s] % kGroup != 0) {
            set_error("scene " + std::to_string(s) + " has " + std::to_string(counts[s]) +
                      " particles, not a non-zero multiple of 256 (cLocalSize)");
            return;
        }
        off_host_[s] = (uint32_t)total;
        total += counts[s];
    }
    if (total >= (1ull << 31)) {
        set_error("batch too large");
        return;
    }
    N_ = (uint32_t)total;
    off_host_[S_] = N_;
    scene_bits_ = 0;
    while ((1u << scene_bits_) < S_) scene_bits_++;
    mfl_params def;
    mfl_params_default(&def);
    params_host_.assign(S_, def);
    box_host_.assign((size_t)S_ * 8, 0.f);
    for (uint32_t s = 0; s < S_; s++) {
        const float mn[4] = {-1.f, -1.f, -1.f, 0.f}, mx[4] = {1.f, 1.f, 1.f, 0.f};
        std::memcpy(&box_host_[(size_t)s * 8], mn, 16);
        std::memcpy(&box_host_[(size_t)s * 8 + 4], mx, 16);
    }
    const size_t v4 = (size_t)N_ * 16, u1 = (size_t)N_ * 4;
    std::vector<uint32_t> pscene(N_);
    for (uint32_t s = 0; s < S_; s++)
        for (uint32_t g = off_host_[s]; g < off_host_[s + 1]; g++) pscene[g] = s;
    const size_t tables = (size_t)S_ * kBuckets * 4;
    if (!off_.allocate((S_ + 1) * 4) || !pscene_.allocate(u1) || !params_.allocate(S_ * sizeof(mfl_params)) ||
        !cuts_.allocate(S_ * sizeof(float)) ||
        !boxes_.allocate((size_t)S_ * 32) || !pos_.allocate(v4) || !vel_.allocate(v4) || !force_.allocate(v4) ||
        !cell_.allocate(u1) || !index_.allocate(u1) || !spos_.allocate(v4) || !svel_.allocate(v4) ||
        !tables_.allocate(2 * tables) || !nbr_.allocate((size_t)kNbrCap * u1) || !ncount_.allocate(u1) ||
        !inv_.allocate(u1) || !porder_.allocate(u1) || !mkey_.allocate(u1))
        return;
    if (!hip_ok(hipMemcpy(off_.ptr, off_host_.data(), (S_ + 1) * 4, hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemcpy(pscene_.ptr, pscene.data(), u1, hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemset(tables_.ptr, 0xFF, 2 * tables), "hipMemset") ||
        !hip_ok(hipMemset(pos_.ptr, 0, v4), "hipMemset") || !hip_ok(hipMemset(vel_.ptr, 0, v4), "hipMemset") ||
        !hip_ok(hipMemset(force_.ptr, 0, v4), "hipMemset") || !hip_ok(hipMemset(spos_.ptr, 0, v4), "hipMemset") ||
        !hip_ok(hipMemset(svel_.ptr, 0, v4), "hipMemset") || !hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize"))
        return;
    ok_ = true;
}

FluidBatch::~FluidBatch()
{
    drop_graph();
    for (DeviceBuffer* b : {&off_, &pscene_, &params_, &boxes_, &pos_, &vel_, &force_, &cell_, &index_, &spos_, &svel_,
                            &tables_, &nbr_, &ncount_, &inv_, &porder_, &mkey_})
        b->release();
}

DeviceBuffer* FluidBatch::select(int kind)
{
    switch (kind) {
    case MFL_POSITION: return &pos_;
    case MFL_VELOCITY: return &vel_;
    case MFL_FORCE: return &force_;
    case MFL_SORTED_POSITION: return &spos_;
    case MFL_SORTED_VELOCITY: return &svel_;
    case MFL_INDEX: return &index_;
    case MFL_CELL: return &cell_;
    case MFL_CELL_START: return &tables_;
    case MFL_CELL_END: return &tables_;
    case MFL_PARAMS: return &params_;
    default: set_error("unknown buffer kind " + std::to_string(kind)); return nullptr;
    }
}

int FluidBatch::get_buffer(int kind, void** ptr, size_t* bytes)
{
    DeviceBuffer* b = select(kind);
    if (!b) return 0;
    const size_t tables = (size_t)S_ * kBuckets * 4;
    char* base = static_cast<char*>(b->ptr);
    if (kind == MFL_CELL_END) base += tables;
    if (ptr) *ptr = base;
    if (bytes) *bytes = (kind == MFL_CELL_START || kind == MFL_CELL_END) ? tables : b->bytes;
    return 1;
}

int FluidBatch::write_buffer(int kind, const void* host, size_t bytes)
{
    void* ptr = nullptr;
    size_t cap = 0;
    if (!get_buffer(kind, &ptr, &cap)) return 0;
    if (bytes > cap) {
        set_error("write of " + std::to_string(bytes) + " bytes exceeds the buffer");
        return 0;
    }
    MFL_HIP(hipMemcpy(ptr, host, bytes, hipMemcpyHostToDevice));
    return 1;
}

int FluidBatch::read_buffer(int kind, void* host, size_t bytes)
{
    void* ptr = nullptr;
    size_t cap = 0;
    if (!get_buffer(kind, &ptr, &cap)) return 0;
    if (bytes > cap) {
        set_error("read of " + std::to_string(bytes) + " bytes exceeds the buffer");
        return 0;
    }
    MFL_HIP(hipDeviceSynchronize());
    MFL_HIP(hipMemcpy(host, ptr, bytes, hipMemcpyDeviceToHost));
    return 1;
}

int FluidBatch::scene_range(uint32_t scene, uint32_t* offset, uint32_t* count) const
{
    if (scene >= S_) {
        set_error("scene index out of range");
        return 0;
    }
    if (offset) *offset = off_host_[scene];
    if (count) *count = off_host_[scene + 1] - off_host_[scene];
    return 1;
}

int FluidBatch::set_params(int scene, const mfl_params& p)
{
    if (scene >= (int)S_) {
        set_error("scene index out of range");
        return 0;
    }
    for (uint32_t s = 0; s < S_; s++)
        if (scene < 0 || (uint32_t)scene == s) params_host_[s] = p;
    params_dirty_ = true;
    return 1;
}

int FluidBatch::get_params(uint32_t scene, mfl_params* p) const
{
    if (scene >= S_ || !p) {
        set_error("scene index out of range");
        return 0;
    }
    *p = params_host_[scene];
    return 1;
}

int FluidBatch::set_box(int scene, const float* mn, const float* mx)
{
    if (scene >= (int)S_ || !mn || !mx) {
        set_error("scene index out of range or null box");
        return 0;
    }
    for (uint32_t s = 0; s < S_; s++)
        if (scene < 0 || (uint32_t)scene == s) {
            std::memcpy(&box_host_[(size_t)s * 8], mn, 16);
            std::memcpy(&box_host_[(size_t)s * 8 + 4], mx, 16);
        }
    params_dirty_ = true;
    return 1;
}

int FluidBatch::enqueue_step(hipStream_t s)
{
    const dim3 b(kThreads), g(grid_for(N_));
    auto* pos = pos_.as<float4>();
    auto* vel = vel_.as<float4>();
    auto* force = force_.as<float4>();
    auto* spos = spos_.as<float4>();
    auto* svel = svel_.as<float4>();
    auto* cell = cell_.as<uint32_t>();
    auto* index = index_.as<uint32_t>();
    auto* cs_all = tables_.as<uint32_t>();
    auto* ce_all = cs_all + (size_t)S_ * kBuckets;
    auto* off = off_.as<uint32_t>();
    auto* params = params_.as<mfl_params>();
    const uint32_t clear = clear_count < kBuckets ? clear_count : kBuckets;

    // the hash kernel also counts the sort's first digit per tile (one count launch fewer)
    SortCountPlan plan;
    if (!sort.first_pass_plan(N_, 0, 18 + (int)scene_bits_, &plan)) return 0;
#define MFL_B_HASH(IT)                                                                                            \
    hipLaunchKernelGGL(k_b_hash_counts<IT>, dim3(plan.ntiles), b, 0, s, pos, pscene_.as<uint32_t>(), params, N_, cell, \
                       plan.counts, plan.ntiles, plan.bits, cs_all, ce_all, clear, S_)
    switch (plan.items) {
    case 16: MFL_B_HASH(16); break;
    case 8: MFL_B_HASH(8); break;
    default: MFL_B_HASH(4); break;
    }
#undef MFL_B_HASH
    MFL_HIP(hipGetLastError());
    if (!sort.sort_bits(cell, index, N_, 0, 18 + (int)scene_bits_, s, true, true)) return 0;
    hipLaunchKernelGGL(k_b_slots_bounds, g, b, 0, s, index, N_, inv_.as<uint32_t>(), cell, off, cs_all, ce_all, pos,
                       vel, spos, svel);
    const bool tiles = (neighbour_pass == MFL_PASS_TILE_LIST || neighbour_pass == MFL_PASS_TILE_KEPT) && porder_valid_;
    const bool kept = tiles && neighbour_pass == MFL_PASS_TILE_KEPT;
    if (kept && !(kept_.ptr && hitmask_.ptr && hitx_.ptr && meta_.ptr)) {
        set_error("kept-entry buffers are not allocated");
        return 0;
    }
    const uint32_t lanes = (uint32_t)density_lanes, ppw = kWave / lanes;
    const size_t waves = ((size_t)N_ + ppw - 1) / ppw;
    size_t blocks = (waves + block_waves - 1) / block_waves;
    blocks = (blocks + kXcds - 1) / kXcds * kXcds;
    if (tiles) {
        const size_t lds = (size_t)block_waves * (8 / lanes) * ((tile_depth + kGroupLanes) * sizeof(float4) +
                                                                sizeof(StageAux));
#define MFL_B_TILE(W, K)                                                                                           \
    hipLaunchKernelGGL((k_b_density_tile<W, K>), dim3((unsigned)blocks), dim3(block_waves * kWave), lds, s, pos, spos, \
                       svel, cs_all, ce_all, off, pscene_.as<uint32_t>(), params, N_, tile_depth,                     \
                       porder_.as<uint32_t>(), inv_.as<uint32_t>(), nbr_.as<uint32_t>(), ncount_.as<uint32_t>(),     \
                       kept_.as<uint32_t>(), hitmask_.as<uint64_t>(), hitx_.as<uint64_t>(), meta_.as<uint32_t>(),   \
                       cuts_.as<float>())
        if (lanes == 2) {
            if (kept) MFL_B_TILE(16, true);
            else MFL_B_TILE(16, false);
        } else if (lanes == 4) {
            if (kept) MFL_B_TILE(32, true);
            else MFL_B_TILE(32, false);
        } else {
            if (kept) MFL_B_TILE(64, true);
            else MFL_B_TILE(64, false);
        }
#undef MFL_B_TILE
    } else {
        hipLaunchKernelGGL(k_b_density, grid_for((size_t)N_ * kGroupLanes), b, 0, s, spos, svel, cs_all, ce_all, cell,
                           off, params, N_, nbr_.as<uint32_t>(), ncount_.as<uint32_t>());
    }
#define MFL_B_FORCES(CH)                                                                                          \
    hipLaunchKernelGGL(k_b_forces_step<CH>, g, b, 0, s, spos, svel, pos, vel, force, index, cell, cs_all, ce_all, off, \
                       params, boxes_.as<float4>(), nbr_.as<uint32_t>(), ncount_.as<uint32_t>(), N_,              \
                       gravity_force_of(gravity))
    if (kept) {
        const size_t lds_k = (size_t)block_waves * (8 / lanes) * (lanes < 8 ? 32 : 64) * 2 * sizeof(float4);
#define MFL_B_KEPT(W)                                                                                              \
    hipLaunchKernelGGL(k_b_forces_kept<W>, dim3((unsigned)blocks), dim3(block_waves * kWave), lds_k, s, spos, svel,  \
                       pos, vel, force, cs_all, ce_all, off, pscene_.as<uint32_t>(), params, boxes_.as<float4>(), N_, \
                       tile_depth, porder_.as<uint32_t>(), inv_.as<uint32_t>(), kept_.as<uint32_t>(), hitmask_.as<uint64_t>(),  \
                       hitx_.as<uint64_t>(), meta_.as<uint32_t>(), gravity_force_of(gravity))
        if (lanes == 2) MFL_B_KEPT(16);
        else if (lanes == 4) MFL_B_KEPT(32);
        else MFL_B_KEPT(64);
#undef MFL_B_KEPT
    } else if (force_chunks == 2) {
        MFL_B_FORCES(2);
    } else {
        MFL_B_FORCES(1);
    }
#undef MFL_B_FORCES
    MFL_HIP(hipGetLastError());
    return 1;
}

void FluidBatch::drop_graph()
{
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_sig_.clear();
}

int FluidBatch::launch_graph(hipStream_t s)
{
    // the sort's temporaries are reserved first: the signature holds their addresses
    if (!sort.reserve(N_, 18 + (int)scene_bits_)) return 0;
    std::string sig;
    sig.reserve(192);
    sort.sig_append(sig);
    sig_add(sig, kept_.ptr);
    sig_add(sig, clear_count);
    sig_add(sig, sort.tuning.items);
    sig_add(sig, gravity);
    sig_add(sig, neighbour_pass);
    sig_add(sig, density_lanes);
    sig_add(sig, tile_depth);
    sig_add(sig, block_waves);
    sig_add(sig, force_chunks);
    sig_add(sig, porder_valid_);
    if (!graph_exec_ || sig != graph_sig_) {
        drop_graph();
        MFL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        const int ok = enqueue_step(s);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &gr);
        if (!ok || e != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            if (ok) hip_ok(e, "hipStreamEndCapture");
            return 0;
        }
        graph_ = gr;
        MFL_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        graph_sig_ = sig;
    }
    MFL_HIP(hipGraphLaunch(graph_exec_, s));
    return 1;
}

bool FluidBatch::fit_kept()
{
    if (kept_.ptr && hitmask_.ptr && hitx_.ptr && meta_.ptr) return true;
    // kept entries of every 8-particle segment, acceptance bits of every lane (kMaxWindows words for a
    // neighbourhood tested in several windows), meta per particle
    const size_t segs = ((size_t)N_ + 63) / 64 * 8;  // every segment of every wave of up to 64 particles
    return kept_.allocate(segs * kSegCap * 4) && hitmask_.allocate((size_t)N_ * 8 * sizeof(uint64_t)) &&
           hitx_.allocate((size_t)N_ * 8 * (kMaxWindows - 1) * sizeof(uint64_t)) && meta_.allocate((size_t)N_ * 4);
}

int FluidBatch::compute(hipStream_t s)
{
    if (!ok_) {
        set_error("batch was not created");
        return 0;
    }
    // the kept-entry pass (the default) writes these; allocate before any launch or graph capture
    if (neighbour_pass == MFL_PASS_TILE_KEPT && !fit_kept()) return 0;
    if (params_dirty_) {
        // the scene Params and boxes are read from device arrays, so their values are not baked into a
        // captured graph; refreshing them needs no re-capture
        std::vector<float4> boxes((size_t)S_ * 2);
        std::memcpy(boxes.data(), box_host_.data(), box_host_.size() * 4);
        MFL_HIP(hipStreamSynchronize(s));  // steps in flight may still read the previous values
        MFL_HIP(hipMemcpy(params_.ptr, params_host_.data(), S_ * sizeof(mfl_params), hipMemcpyHostToDevice));
        std::vector<float> cuts(S_);
        for (uint32_t q = 0; q < S_; q++) cuts[q] = sqrt_cut(params_host_[q].cellSize);
        MFL_HIP(hipMemcpy(cuts_.ptr, cuts.data(), S_ * sizeof(float), hipMemcpyHostToDevice));
        MFL_HIP(hipMemcpy(boxes_.ptr, boxes.data(), boxes.size() * 16, hipMemcpyHostToDevice));
        params_dirty_ = false;
    }
    if (neighbour_pass != MFL_PASS_LIST &&
        (!porder_valid_ || (refresh_interval > 0 && steps_since_refresh_ >= refresh_interval))) {
        if (!refresh_order(s)) return 0;
        steps_since_refresh_ = 0;
    }
    steps_since_refresh_++;
    if (use_graph && s != nullptr) return launch_graph(s);
    return enqueue_step(s);
}

int FluidBatch::refresh_order(hipStream_t s)
{
    const uint32_t mbits = (32u - scene_bits_) / 3u > 10u ? 10u : (32u - scene_bits_) / 3u;
    hipLaunchKernelGGL(k_b_morton_keys, grid_for(N_), dim3(kThreads), 0, s, pos_.as<float4>(), pscene_.as<uint32_t>(),
                       params_.as<mfl_params>(), N_, mbits, mkey_.as<uint32_t>());
    MFL_HIP(hipGetLastError());
    if (!sort.sort_bits(mkey_.as<uint32_t>(), porder_.as<uint32_t>(), N_, 0, (int)(3 * mbits + scene_bits_), s, true))
        return 0;
    porder_valid_ = true;
    return 1;
}

int FluidBatch::get_tuning(mfl_tuning* out) const
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

int FluidBatch::set_tuning(const mfl_tuning& t)
{
    if (t.neighbour_pass != MFL_PASS_LIST && t.neighbour_pass != MFL_PASS_TILE_LIST &&
        t.neighbour_pass != MFL_PASS_TILE_KEPT) {
        set_error("batched scenes take MFL_PASS_LIST, MFL_PASS_TILE_LIST or MFL_PASS_TILE_KEPT");
        return 0;
    }
    const bool tile = t.neighbour_pass != MFL_PASS_LIST;
    if (tile ? (t.density_lanes != 2 && t.density_lanes != 4 && t.density_lanes != 8) : t.density_lanes != 8) {
        set_error(tile ? "density_lanes must be 2, 4 or 8 with the tile passes"
                       : "density_lanes must be 8 with MFL_PASS_LIST for batched scenes");
        return 0;
    }
    if (t.tile_depth < kTileMinDepth || t.tile_depth > kTileMaxDepth || t.tile_depth % 16 != 0 ||
        (tile && t.tile_depth > 64u * (uint32_t)t.density_lanes)) {
        set_error("tile_depth must be a multiple of 16 in [64, 64 x density_lanes]");
        return 0;
    }
    if (t.block_waves != 1 && t.block_waves != 2 && t.block_waves != 4) {
        set_error("block_waves must be 1, 2 or 4");
        return 0;
    }
    if (tile && (size_t)t.block_waves * (8 / t.density_lanes) *
                        ((t.tile_depth + kGroupLanes) * sizeof(float4) + sizeof(StageAux)) > 65536) {
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
    if (t.neighbour_pass == MFL_PASS_TILE_KEPT && !fit_kept()) return 0;
    neighbour_pass = t.neighbour_pass;
    density_lanes = t.density_lanes;
    tile_depth = t.tile_depth;
    block_waves = t.block_waves;
    sort.tuning.items = (int)t.sort_items;
    force_chunks = t.force_chunks;
    return 1;
}

int FluidBatch::autotune(hipStream_t s, uint32_t steps, mfl_tuning* best, double* ms_per_step)
{
    if (!ok_) {
        set_error("batch was not created");
        return 0;
    }
    if (steps == 0) {
        set_error("autotune needs at least one step per trial");
        return 0;
    }
    // the trials re-size the sort's temporaries (other tile sizes): no graph may keep the old ones
    drop_graph();
    const size_t vb = (size_t)N_ * sizeof(float4), tb = (size_t)S_ * kBuckets * 2 * sizeof(uint32_t);
    DeviceBuffer save;
    if (!save.allocate(3 * vb + tb)) return 0;
    char* sp = save.as<char>();
    auto snapshot = [&](bool restore) -> int {
        void* dev[] = {pos_.ptr, vel_.ptr, force_.ptr, tables_.ptr};
        const size_t len[] = {vb, vb, vb, tb};
        size_t o = 0;
        for (int k = 0; k < 4; k++) {
            void* dst = restore ? dev[k] : sp + o;
            const void* src = restore ? (const void*)(sp + o) : dev[k];
            MFL_HIP(hipMemcpyAsync(dst, src, len[k], hipMemcpyDeviceToDevice, s));
            o += len[k];
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
    cand.push_back({MFL_PASS_LIST, 8, 192u, 2u, start.sort_items, 2u});
    for (uint32_t w : {1u, 2u, 4u})
        for (uint32_t d : {96u, 128u, 192u, 256u, 384u})
            for (int lanes : {2, 4, 8})
                for (int pass : {MFL_PASS_TILE_LIST, MFL_PASS_TILE_KEPT})
                    for (uint32_t ch : {1u, 2u}) {
                        if (pass == MFL_PASS_TILE_KEPT && ch == 2u) continue;  // chunks only shape the list pass
                        const mfl_tuning t{pass, lanes, d, w, start.sort_items, ch};
                        if (d <= 64u * (uint32_t)lanes &&
                            (size_t)w * (8 / lanes) * ((d + kGroupLanes) * sizeof(float4) + sizeof(StageAux)) <= 65536)
                            cand.push_back(t);
                    }
    const bool graph_was = use_graph;
    const uint32_t since = steps_since_refresh_;
    use_graph = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int ok = hip_ok(hipEventCreate(&e0), "hipEventCreate") && hip_ok(hipEventCreate(&e1), "hipEventCreate");
    // The card lowers its clocks within milliseconds of idling: step the starting shape untimed for 30 ms
    // first, so that the first candidates are not timed at lower clocks than the last ones.
    {
        const auto t0 = std::chrono::steady_clock::now();
        while (ok && std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(30)) {
            ok = snapshot(true);
            for (int j = 0; ok && j < 4; j++) ok = compute(s);
            ok = ok && hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
        }
    }
    double best_ms = 1e30;
    mfl_tuning pick = start;
    for (size_t c = 0; ok && c < cand.size(); c++) {
        ok = set_tuning(cand[c]) && snapshot(true) && compute(s);  // one untimed step per trial
        for (uint32_t k = 0; ok && k < steps; k++) {
            if (k == 0) ok = hip_ok(hipEventRecord(e0, s), "hipEventRecord");
            ok = ok && compute(s);
        }
        ok = ok && hip_ok(hipEventRecord(e1, s), "hipEventRecord") && hip_ok(hipEventSynchronize(e1), "sync");
        float ms = 0.f;
        ok = ok && hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (ok && ms / steps < best_ms) {
            best_ms = ms / steps;
            pick = cand[c];
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    use_graph = graph_was;
    const int restored = snapshot(true) && hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
    steps_since_refresh_ = since;
    porder_valid_ = false;  // rebuilt from the restored positions on the next step
    set_tuning(ok ? pick : start);
    save.release();
    if (!ok || !restored) return 0;
    if (best) *best = pick;
    if (ms_per_step) *ms_per_step = best_ms;
    return 1;
}

}  // namespace mfl
