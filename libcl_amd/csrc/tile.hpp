// LDS-staged neighbourhoods for the density and force passes (clCalculateDensity / clCalculateForces,
// libCL/phys/oclFluid3D.cl:150-255).
//
// A wave holds the particles of 8 consecutive traversal slots, 8 lanes each. Particles that share a
// cell share the 27 bucket ranges the reference walks for them, so the wave stages those ranges once:
// lanes 0..26 read one range each (dx fastest, then dy, then dz, as cl:166-170); a prefix sum lays the
// ranges end to end (up to 512 visits), and the wave loads them 64 at a time, keeping, in visit order,
// every entry that can lie within h of the box spanned by the wave's particles of that cell. The
// particles then test only the kept entries, from LDS. Pruning is exact: an entry is dropped only when
// its distance to the box exceeds h (1 + 5e-5), where the reference's fast_length(d) < h test is false
// for every particle in the box. Repeated buckets are staged once per visit and stale ranges as they
// stand, so every particle sees the reference's visit multiset.
#pragma once

#include "sph_device.hpp"

namespace mfl {
namespace sph {

constexpr float kKeepMargin = 1.0001f;      // h^2 (1 + 1e-4): squared-distance margin of the box prune
constexpr uint32_t kStageRounds = 8;        // visits of one neighbourhood laid end to end: at most 8 x 64
constexpr uint32_t kTileMinDepth = 64, kTileMaxDepth = kStageRounds * 64;  // kept <= staged <= 512
// Kept-entry passes: the density pass writes each 8-particle segment's kept entries (at most kSegCap per
// segment over its cells) and, per particle, meta = offset | kept << 16 of its cell's entries, or
// kMetaWalk when the force pass must walk the bucket ranges itself; the force pass stages kept_chunk<W>
// entries per segment at a time: 32 when a wave holds two segments or more (one 64-lane state load then
// fills two segments' chunks), which halves a wave's LDS and so lets more waves hide the loads' latency;
// 64 for a one-segment wave.
constexpr uint32_t kSegCap = 512;
constexpr uint32_t kMetaWalk = 0xFFFFFFFFu;
template <int W>
constexpr uint32_t kept_chunk()
{
    return W < kWave ? 32u : 64u;
}

// Per-wave LDS bookkeeping of a staging: bit p % 64 of starts[p / 64] marks position p where a
// non-empty bucket range begins, and delta[c] maps a position of the c-th non-empty range to its
// sorted index (j = p + delta[c]). The segmented staging also keeps the cell's box (lo xyz, hi xyz) and
// its number of visits there, for the windows staged after the first.
struct StageAux {
    unsigned long long starts[kStageRounds];
    uint32_t delta[32];
    float box[6];
    uint32_t total;
};

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ inline float group_sum(float v)
{
    v += __shfl_xor(v, 1, kWave);
    v += __shfl_xor(v, 2, kWave);
    v += __shfl_xor(v, 4, kWave);
    return v;
}

__device__ inline uint32_t group_sum_u(uint32_t v)
{
    v += (uint32_t)__shfl_xor((int)v, 1, kWave);
    v += (uint32_t)__shfl_xor((int)v, 2, kWave);
    v += (uint32_t)__shfl_xor((int)v, 4, kWave);
    return v;
}

// Rounds r0 .. r0 + NR - 1 of a staging (positions 64 r + lane): loads, box prune and in-order
// compaction of the kept entries into the tile.
template <int NR, bool FORCES>
__device__ inline void stage_rounds(uint32_t r0, uint32_t total, uint32_t lane, uint64_t upto, float3 lo, float3 hi,
                                    float h2_keep, const float4* __restrict__ spos, uint32_t depth, float4* tpos,
                                    uint32_t* tj, const StageAux* aux, uint32_t& kept, uint32_t& before)
{
    float4 v[NR];
    uint32_t jv[NR];
    bool has[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const uint32_t q = (r0 + r) * kWave + lane;
        const unsigned long long sm = aux->starts[r0 + r];
        has[r] = q < total;
        const uint32_t c = before + (uint32_t)__popcll(sm & upto) - 1u;  // non-empty ranges starting at or before q
        before += (uint32_t)__popcll(sm);
        jv[r] = q + aux->delta[c & 31u];
        if (has[r]) v[r] = spos[jv[r]];
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        bool keep = false;
        if (has[r]) {
            const float ex = fmaxf(fmaxf(lo.x - v[r].x, v[r].x - hi.x), 0.f);
            const float ey = fmaxf(fmaxf(lo.y - v[r].y, v[r].y - hi.y), 0.f);
            const float ez = fmaxf(fmaxf(lo.z - v[r].z, v[r].z - hi.z), 0.f);
            keep = dist2(ex, ey, ez) < h2_keep;
        }
        const uint64_t m = __ballot(keep);
        if (keep) {
            const uint32_t at = kept + lanes_below(m);
            if (at < depth) {
                if (FORCES) {
                    tpos[at] = v[r];
                    tj[at] = jv[r];
                } else {
                    tpos[at] = make_float4(v[r].x, v[r].y, v[r].z, __uint_as_float(jv[r]));
                }
            }
        }
        kept += (uint32_t)__popcll(m);
    }
}

// Stages the visits of cell (cx, cy, cz)'s 27 bucket ranges that can lie within h of [lo, hi] and
// returns how many there are (at most `depth` are written), or kStageOverflow when the ranges hold more
// than kStageRounds x 64 visits. FORCES = false writes float4(x, y, z, bits of j) to tpos; FORCES = true
// writes spos[j] to tpos, svel[j] to tvel and j to tj. Called by the whole wave with uniform arguments.
constexpr uint32_t kStageOverflow = 0xFFFFFFFFu;

template <bool FORCES>
__device__ inline uint32_t stage_neighbourhood(int cx, int cy, int cz, float3 lo, float3 hi, float h2_keep,
                                               const float4* __restrict__ spos, const float4* __restrict__ svel,
                                               const uint32_t* __restrict__ cell_start,
                                               const uint32_t* __restrict__ cell_end, uint32_t n, uint32_t lane,
                                               uint32_t depth, float4* tpos, float4* tvel, uint32_t* tj,
                                               StageAux* aux)
{
    uint32_t st = 0, len = 0;
    if (lane < 27) {
        const int dx = (int)(lane % 3) - 1, dy = (int)((lane / 3) % 3) - 1, dz = (int)(lane / 9) - 1;
        const uint32_t c = cell_hash(cx + dx, cy + dy, cz + dz);
        const uint32_t s = cell_start[c];
        const uint32_t e = min(cell_end[c], n);  // stale or written ranges always lie in [0, n]
        if (s != kEmpty) {
            st = s;
            len = e > s ? e - s : 0u;
        }
    }
    // ranges laid end to end in visit order: range k starts at position pre
    uint32_t incl = len;
#pragma unroll
    for (uint32_t o = 1; o < 32; o <<= 1) {
        const uint32_t up = (uint32_t)__shfl_up((int)incl, o, kWave);
        if ((lane & 31u) >= o) incl += up;
    }
    const uint32_t pre = incl - len;
    const uint32_t total = __builtin_amdgcn_readlane(incl, 26);
    if (total > kStageRounds * kWave) return kStageOverflow;
    if (lane < kStageRounds) aux->starts[lane] = 0ull;
    wave_sync();
    const uint64_t nonempty = __ballot(len > 0);
    if (len > 0) {
        aux->delta[lanes_below(nonempty)] = st - pre;
        atomicOr(&aux->starts[pre / kWave], 1ull << (pre % kWave));
    }
    wave_sync();
    const uint64_t upto = lane == kWave - 1 ? ~0ull : (2ull << lane) - 1ull;  // bits 0..lane
    const uint32_t rounds = (total + kWave - 1) / kWave;
    uint32_t kept = 0, before = 0, r0 = 0;
    // rounds of 64 positions go four (then two, then one) at a time, so that their loads are in flight
    // together without staging rounds past the end
    for (; r0 + 4 <= rounds; r0 += 4)
        stage_rounds<4, FORCES>(r0, total, lane, upto, lo, hi, h2_keep, spos, depth, tpos, tj, aux, kept, before);
    if (r0 + 2 <= rounds) {
        stage_rounds<2, FORCES>(r0, total, lane, upto, lo, hi, h2_keep, spos, depth, tpos, tj, aux, kept, before);
        r0 += 2;
    }
    if (r0 < rounds)
        stage_rounds<1, FORCES>(r0, total, lane, upto, lo, hi, h2_keep, spos, depth, tpos, tj, aux, kept, before);
    if (FORCES) {
        wave_sync();
        const uint32_t m = min(kept, depth);
        for (uint32_t q = lane; q < m; q += kWave) tvel[q] = svel[tj[q]];
    }
    wave_sync();
    return kept;
}

// Min / max over the eight particle groups of a wave (the lanes of a group hold the same value).
__device__ inline float groups_min(float v)
{
    v = fminf(v, __shfl_xor(v, 8, kWave));
    v = fminf(v, __shfl_xor(v, 16, kWave));
    return fminf(v, __shfl_xor(v, 32, kWave));
}

__device__ inline float groups_max(float v)
{
    v = fmaxf(v, __shfl_xor(v, 8, kWave));
    v = fmaxf(v, __shfl_xor(v, 16, kWave));
    return fmaxf(v, __shfl_xor(v, 32, kWave));
}

// The cell of the first remaining lane, whether this lane's particle is in it, and the bounding box
// of the particles that are.
struct CellGroup {
    int cx, cy, cz;
    bool member;
    float3 lo, hi;
};

__device__ inline CellGroup next_cell(uint64_t todo, bool live, int gx, int gy, int gz, const float4& pi)
{
    CellGroup c;
    const int src = __ffsll((unsigned long long)todo) - 1;
    c.cx = __builtin_amdgcn_readlane(gx, src);
    c.cy = __builtin_amdgcn_readlane(gy, src);
    c.cz = __builtin_amdgcn_readlane(gz, src);
    c.member = live && gx == c.cx && gy == c.cy && gz == c.cz;
    const float inf = __builtin_inff();
    c.lo = make_float3(groups_min(c.member ? pi.x : inf), groups_min(c.member ? pi.y : inf),
                       groups_min(c.member ? pi.z : inf));
    c.hi = make_float3(groups_max(c.member ? pi.x : -inf), groups_max(c.member ? pi.y : -inf),
                       groups_max(c.member ? pi.z : -inf));
    return c;
}

// ---- segmented tiles: W-lane segments of a wave, each staging the cell of its own 8 particles ----------
//
// With W = 32 a wave holds 16 particles (4 lanes each) of up to two cells and stages both
// neighbourhoods at once, halving the per-cell cost of everything that runs once per staging (range
// fetch, prefix sum, cell and box selection, reductions). W = 64 is the one-cell-per-wave layout.

template <int W>
struct Lanes {
    static constexpr uint32_t kLpp = W / 8;  // lanes per particle
    uint32_t lane, seg, ls;                   // lane of the wave, segment, lane within the segment
    __device__ explicit Lanes(uint32_t l) : lane(l), seg(l / W), ls(l % W) {}
    // this segment's bits of a wave ballot, as bits 0 .. W-1
    __device__ uint64_t bits(uint64_t b) const
    {
        if constexpr (W == 64) return b;
        else return (b >> (seg * W)) & ((1ull << W) - 1ull);
    }
    __device__ uint32_t below(uint64_t segbits) const { return (uint32_t)__popcll(segbits & ((1ull << ls) - 1ull)); }
    __device__ int from(int v, uint32_t k) const { return __shfl(v, (int)(seg * W + k), kWave); }
    __device__ uint32_t from(uint32_t v, uint32_t k) const { return (uint32_t)__shfl((int)v, (int)(seg * W + k), kWave); }
};

// min / max / sum over the 8 particles of a segment (the lanes of one particle hold equal values)
template <int W>
__device__ inline float seg_groups_min(float v)
{
    constexpr int P = W / 8;
    v = fminf(v, __shfl_xor(v, P, kWave));
    v = fminf(v, __shfl_xor(v, 2 * P, kWave));
    return fminf(v, __shfl_xor(v, 4 * P, kWave));
}

template <int W>
__device__ inline float seg_groups_max(float v)
{
    constexpr int P = W / 8;
    v = fmaxf(v, __shfl_xor(v, P, kWave));
    v = fmaxf(v, __shfl_xor(v, 2 * P, kWave));
    return fmaxf(v, __shfl_xor(v, 4 * P, kWave));
}

// sum over the lanes of one particle
template <int W>
__device__ inline float particle_sum(float v)
{
#pragma unroll
    for (int o = 1; o < W / 8; o <<= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

template <int W>
__device__ inline uint32_t particle_sum_u(uint32_t v)
{
#pragma unroll
    for (int o = 1; o < W / 8; o <<= 1) v += (uint32_t)__shfl_xor((int)v, o, kWave);
    return v;
}

struct SegCell {
    int cx, cy, cz;
    bool active;  // the segment has a cell this round
    bool member;  // this lane's particle is in it
    float3 lo, hi;
};

// The cell of the segment's first remaining particle (todo: the segment's W-bit mask of lanes still to
// serve), whether this lane's particle is in it, and the box of the particles that are.
template <int W>
__device__ inline SegCell next_cell_seg(const Lanes<W>& L, uint64_t todo, bool live, int gx, int gy, int gz,
                                        const float4& pi)
{
    SegCell c;
    c.active = todo != 0;
    const uint32_t src = c.active ? (uint32_t)(__ffsll((unsigned long long)todo) - 1) : 0u;
    c.cx = L.from(gx, src);
    c.cy = L.from(gy, src);
    c.cz = L.from(gz, src);
    c.member = c.active && live && gx == c.cx && gy == c.cy && gz == c.cz;
    const float inf = __builtin_inff();
    c.lo = make_float3(seg_groups_min<W>(c.member ? pi.x : inf), seg_groups_min<W>(c.member ? pi.y : inf),
                       seg_groups_min<W>(c.member ? pi.z : inf));
    c.hi = make_float3(seg_groups_max<W>(c.member ? pi.x : -inf), seg_groups_max<W>(c.member ? pi.y : -inf),
                       seg_groups_max<W>(c.member ? pi.z : -inf));
    return c;
}

// Rounds r0 .. r0 + NR - 1 of a segment's staging (positions W r + ls): loads, box prune, in-order
// compaction of the kept entries into the tile as float4(x, y, z, bits of j). Kept entry k goes to tile
// slot k - skip when skip <= k < skip + depth (one window of the kept list).
template <int W, int NR>
__device__ inline void stage_seg_rounds(const Lanes<W>& L, uint32_t r0, uint32_t total, uint64_t upto, float3 lo,
                                        float3 hi, float h2_keep, const float4* __restrict__ spos, uint32_t skip,
                                        uint32_t depth, float4* tile, const StageAux* aux, uint32_t& kept,
                                        uint32_t& before)
{
    float4 v[NR];
    uint32_t jv[NR], cr[NR];
    bool has[NR];
    // phases, so that the LDS reads of all NR rounds and then their global loads are in flight together:
    // start masks, range index of each position, range offsets, positions
    unsigned long long sm[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) sm[r] = aux->starts[((r0 + r) * W) / 64];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        if constexpr (W < 64) sm[r] = (sm[r] >> (((r0 + r) * W) % 64)) & ((1ull << W) - 1ull);
        cr[r] = before + (uint32_t)__popcll(sm[r] & upto) - 1u;  // non-empty ranges starting at or before q
        before += (uint32_t)__popcll(sm[r]);
    }
#pragma unroll
    for (int r = 0; r < NR; r++) cr[r] = aux->delta[cr[r] & 31u];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const uint32_t q = (r0 + r) * W + L.ls;
        has[r] = q < total;
        jv[r] = q + cr[r];
        v[r] = spos[has[r] ? jv[r] : 0u];  // lanes past the end load entry 0 (n >= 1) and keep nothing
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        // offset of the entry from the box (v minus its clamp to [lo, hi]); NaN positions are dropped,
        // as no particle accepts them
        const float ex = v[r].x - __builtin_amdgcn_fmed3f(v[r].x, lo.x, hi.x);
        const float ey = v[r].y - __builtin_amdgcn_fmed3f(v[r].y, lo.y, hi.y);
        const float ez = v[r].z - __builtin_amdgcn_fmed3f(v[r].z, lo.z, hi.z);
        const bool keep = (dist2(ex, ey, ez) < h2_keep) & has[r];
        const uint64_t m = L.bits(__ballot(keep));
        if (keep) {
            const uint32_t at = kept + L.below(m) - skip;  // wraps above depth when below the window
            if (at < depth) tile[at] = make_float4(v[r].x, v[r].y, v[r].z, __uint_as_float(jv[r]));
        }
        kept += (uint32_t)__popcll(m);
    }
}

// The largest value of v over the wave's segments (v equal within each segment), wave-uniform.
template <int W>
__device__ inline uint32_t seg_wave_max(uint32_t v)
{
#pragma unroll
    for (int o = W; o < kWave; o <<= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, kWave));
    return __builtin_amdgcn_readfirstlane(v);
}

// Segmented stage_neighbourhood (density form), first half: lays the 27 bucket ranges of the segment's
// cell end to end in `aux` (a segment with active = false lays none) and returns the number of visits,
// 0 for an inactive segment, or kStageOverflow when they exceed kStageRounds x 64. A segment of W lanes
// holds the 27 ranges in ceil(27 / W) layers: bucket k sits in lane k % W of layer k / W.
template <int W>
__device__ inline uint32_t stage_seg_ranges(const Lanes<W>& L, const SegCell& c, const uint32_t* __restrict__ cell_start,
                                            const uint32_t* __restrict__ cell_end, uint32_t n, StageAux* aux)
{
    constexpr int NB = (27 + W - 1) / W;  // layers
    constexpr uint32_t SW = W < 32 ? W : 32;  // scan width: the 27 ranges of a wider segment are in its first 32 lanes
    uint32_t st[NB], len[NB], pre[NB];
    uint32_t total = 0;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const uint32_t k = L.ls + (uint32_t)b * W;
        st[b] = 0;
        len[b] = 0;
        if (c.active && k < 27) {
            const int dx = (int)(k % 3) - 1, dy = (int)((k / 3) % 3) - 1, dz = (int)(k / 9) - 1;
            const uint32_t bk = cell_hash(c.cx + dx, c.cy + dy, c.cz + dz);
            const uint32_t s = cell_start[bk];
            const uint32_t e = min(cell_end[bk], n);  // stale or written ranges always lie in [0, n]
            if (s != kEmpty) {
                st[b] = s;
                len[b] = e > s ? e - s : 0u;
            }
        }
        // ranges laid end to end in visit order: layer b after every range of the earlier layers
        uint32_t incl = len[b];
#pragma unroll
        for (uint32_t o = 1; o < SW; o <<= 1) {
            const uint32_t up = (uint32_t)__shfl_up((int)incl, o, SW);
            if ((L.lane & (SW - 1u)) >= o) incl += up;
        }
        pre[b] = total + incl - len[b];
        total += L.from(incl, (uint32_t)min(26 - b * W, W - 1));
    }
    const bool over = total > kStageRounds * kWave;
    if (L.ls < kStageRounds) aux->starts[L.ls] = 0ull;
    wave_sync();
    uint32_t ranks = 0;  // non-empty ranges of the earlier layers
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const bool mark = len[b] > 0 && !over;
        const uint64_t nonempty = L.bits(__ballot(mark));
        if (mark) {
            aux->delta[ranks + L.below(nonempty)] = st[b] - pre[b];
            atomicOr(&aux->starts[pre[b] / 64], 1ull << (pre[b] % 64));
        }
        ranks += (uint32_t)__popcll(nonempty);
    }
    if (L.ls == 0) {
        aux->box[0] = c.lo.x;
        aux->box[1] = c.lo.y;
        aux->box[2] = c.lo.z;
        aux->box[3] = c.hi.x;
        aux->box[4] = c.hi.y;
        aux->box[5] = c.hi.z;
        aux->total = over ? 0u : total;
    }
    wave_sync();
    return over ? kStageOverflow : total;
}

// Second half: loads the `total` laid-out visits (0 stages nothing), keeps those that can lie within h
// of the cell's box [c.lo, c.hi] and writes kept entries skip .. skip + depth - 1 to the tile. Returns
// the segment's kept count (over the whole list, whatever the window). Called by the whole wave.
template <int W>
__device__ inline uint32_t stage_seg_fill(const Lanes<W>& L, uint32_t total, float3 lo, float3 hi, float h2_keep,
                                          const float4* __restrict__ spos, uint32_t skip, uint32_t depth, float4* tile,
                                          const StageAux* aux)
{
    const uint32_t rounds = seg_wave_max<W>((total + W - 1) / W);  // the wave runs the longest segment's rounds
    const uint64_t upto = L.ls == 63 ? ~0ull : (2ull << L.ls) - 1ull;  // bits 0..ls
    uint32_t kept = 0, before = 0, r0 = 0;
    for (; r0 + 4 <= rounds; r0 += 4)
        stage_seg_rounds<W, 4>(L, r0, total, upto, lo, hi, h2_keep, spos, skip, depth, tile, aux, kept, before);
    if (r0 + 2 <= rounds) {
        stage_seg_rounds<W, 2>(L, r0, total, upto, lo, hi, h2_keep, spos, skip, depth, tile, aux, kept, before);
        r0 += 2;
    }
    if (r0 < rounds) stage_seg_rounds<W, 1>(L, r0, total, upto, lo, hi, h2_keep, spos, skip, depth, tile, aux, kept, before);
    wave_sync();
    return kept;
}

// Both halves: the segment's kept count, its first `depth` kept entries in the tile, or kStageOverflow.
template <int W>
__device__ inline uint32_t stage_seg(const Lanes<W>& L, const SegCell& c, float h2_keep, const float4* __restrict__ spos,
                                     const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end,
                                     uint32_t n, uint32_t depth, float4* tile, StageAux* aux)
{
    const uint32_t total = stage_seg_ranges<W>(L, c, cell_start, cell_end, n, aux);
    const uint32_t kept =
        stage_seg_fill<W>(L, total == kStageOverflow ? 0u : total, c.lo, c.hi, h2_keep, spos, 0u, depth, tile, aux);
    return total == kStageOverflow ? kStageOverflow : kept;
}

// Acceptance bits of one chunk [c0, c0 + CH) of a window's kept entries (c0 a multiple of CH), from a
// lane's word hm over the window's `trips` trips (bit trips - 1 - m is trip m): trip c0 / P + k of the
// chunk becomes bit 31 - k of the result, so counting leading zeros visits the accepted entries of the
// chunk earliest trip first, entry k P + slot of the chunk for k leading zeros.
template <uint32_t P, uint32_t CH>
__device__ inline uint32_t chunk_bits(uint64_t hm, uint32_t trips, uint32_t c0)
{
    static_assert(CH % P == 0 && CH / P <= 32, "a chunk's trips fit one word");
    const uint32_t mlo = c0 / P, mhi = min((c0 + CH) / P, trips);
    if (mlo >= mhi) return 0u;
    const uint32_t blo = trips - mhi, nb = mhi - mlo;
    return (uint32_t)((hm >> blo) << (32u - nb));  // bits of earlier chunks leave the word
}

// Kept-entry force pass: stages entries c0 .. c0 + CH - 1 (CH = kept_chunk<W>) of the current window of
// every segment of the wave into the wave's LDS tiles, segment e's entries at wpos / wvel[e CH ..]: each
// 64-lane load covers 64 / CH segments (lane l: segment l / CH of its group, entry c0 + l % CH), the
// lanes read consecutive indices of their segment's kept list (the window starts at kept_in[kbase] and
// holds nw entries, both per segment) and the states they name go straight from global memory to LDS
// (global_load_lds_dwordx4: lane l fills slot l of the 1 KiB block given), so no state passes through
// registers. Lanes past a window load spos / svel[0] into slots never used.
// In two halves, so that a chunk's index loads can be in flight while the chunk before it is swept:
// kept_chunk_index issues the index loads of chunk c0 (raw values, used by kept_chunk_load only),
// kept_chunk_load turns them into state loads into LDS and returns without waiting for those.
template <int W>
__device__ inline void kept_chunk_index(const uint32_t* __restrict__ kept_in, size_t kbase, uint32_t nw, uint32_t c0,
                                        uint32_t lane, uint32_t* k)
{
    constexpr uint32_t SEGS = kWave / W, CH = kept_chunk<W>(), PER = kWave / CH, LOADS = SEGS / PER;
    static_assert(SEGS % PER == 0, "whole segments per load");
    const uint32_t half = lane / CH, q = c0 + lane % CH;
#pragma unroll
    for (uint32_t i = 0; i < LOADS; i++) {
        size_t b = 0;
        uint32_t m = 0;
#pragma unroll
        for (uint32_t h = 0; h < PER; h++) {  // this lane's segment: i PER + half
            const uint32_t e = i * PER + h;
            const size_t be = ((size_t)__builtin_amdgcn_readlane((uint32_t)(kbase >> 32), e * W) << 32) |
                              __builtin_amdgcn_readlane((uint32_t)kbase, e * W);
            const uint32_t me = __builtin_amdgcn_readlane(nw, e * W);
            b = h == half ? be : b;
            m = h == half ? me : m;
        }
        k[i] = kept_in[b + min(q, (m > 0 ? m : 1u) - 1u)];  // inside the segment's list
    }
}

template <int W>
__device__ inline void kept_chunk_load(const uint32_t* k, uint32_t nw, uint32_t c0, uint32_t lane,
                                       const float4* __restrict__ spos, const float4* __restrict__ svel, float4* wpos,
                                       float4* wvel)
{
    constexpr uint32_t SEGS = kWave / W, CH = kept_chunk<W>(), PER = kWave / CH, LOADS = SEGS / PER;
    const uint32_t half = lane / CH, q = c0 + lane % CH;
    uint32_t j[LOADS];
    // every index before the first state load (s_waitcnt vmcnt(0) alone: expcnt 7, lgkmcnt 15)
    __builtin_amdgcn_s_waitcnt(0x0F70);
#pragma unroll
    for (uint32_t i = 0; i < LOADS; i++) {
        uint32_t m = 0;
#pragma unroll
        for (uint32_t h = 0; h < PER; h++) {
            const uint32_t me = __builtin_amdgcn_readlane(nw, (i * PER + h) * W);
            m = h == half ? me : m;
        }
        j[i] = q < m ? k[i] : 0u;
    }
#if defined(__HIP_DEVICE_COMPILE__)  // the LDS address space exists in the device pass only
#pragma unroll
    for (uint32_t i = 0; i < LOADS; i++) {
        __builtin_amdgcn_global_load_lds(spos + j[i], (__attribute__((address_space(3))) void*)(wpos + i * kWave),
                                         16, 0, 0);
        __builtin_amdgcn_global_load_lds(svel + j[i], (__attribute__((address_space(3))) void*)(wvel + i * kWave),
                                         16, 0, 0);
    }
#endif
}

// Trip count of a staged neighbourhood of `kept` entries for the wave: a lane tests entries P m + slot
// in trips m = 0, 1, ...; a wave runs the longest of its segments' trip counts.
template <int W>
__device__ inline uint32_t wave_trips(uint32_t kept)
{
    constexpr uint32_t P = W / 8;
    return seg_wave_max<W>((kept + P - 1) / P);
}

// End markers (positions at infinity, index kEmpty) in entries [kept, end) of a segment's tile, so that
// every entry a lane reads past its segment's end is rejected without a bound test. With end = wtrips P
// taken over segments whose kept <= depth, end <= depth + P - 1 < depth + 8, the tile's size.
template <int W>
__device__ inline void pad_tile(const Lanes<W>& L, float4* tile, uint32_t kept, uint32_t end)
{
    const float inf = __builtin_inff();
    for (uint32_t q = kept + L.ls; q < end; q += W) tile[q] = make_float4(inf, inf, inf, __uint_as_float(kEmpty));
}

// One lane's acceptance sweep over a staged neighbourhood: trip m (of the wave's wtrips <= 64) tests
// entry P m of `row` (= tile + slot). An entry is accepted when it is not the particle itself and its
// squared distance is below `cut` (sqrt(r2) < h, exactly). The accepted (h^2 - r2)^3 terms are added to
// `a` in trip order and bit wtrips - 1 - m of the result records trip m. The padded tile lets the
// reads use one base address and immediate offsets.
template <uint32_t P>
__device__ inline uint64_t accept_sweep(const float4* __restrict__ row, uint32_t wtrips, const float4& pi,
                                        uint32_t self, float h2, float cut, float& a)
{
    auto test = [&](uint32_t m) -> uint32_t {
        const float4 e = row[m * P];
        const float dx = pi.x - e.x, dy = pi.y - e.y, dz = pi.z - e.z;
        const float r2 = dist2(dx, dy, dz);
        const bool ok = __float_as_uint(e.w) != self && r2 < cut;
        const float w = h2 - r2;  // h^2 - vlen^2 of cl:123
        a = ok ? fmaf(w * w, w, a) : a;
        return ok ? 1u : 0u;
    };
    uint64_t hits = 0;
#pragma unroll 4
    for (uint32_t m = 0; m < wtrips; m++) hits = (hits << 1) + test(m);
    return hits;
}

// Windows of a kept list: a staged cell's kept entries are tested `depth` at a time from LDS, window w
// holding entries w depth .. w depth + depth - 1, each window one 64-bit word of acceptance bits per
// lane (depth <= 64 x lanes per particle). Window 0 goes to the per-slot hitmask array, windows 1 ..
// kMaxWindows - 1 to an array of kMaxWindows - 1 words per slot; a cell with more windows is walked.
constexpr uint32_t kMaxWindows = 4;

__device__ inline size_t hitx_at(uint32_t t, uint32_t lanes, uint32_t slot, uint32_t w)
{
    return ((size_t)t * lanes + slot) * (kMaxWindows - 1) + (w - 1);
}

// Sweeps the segment's particles over every window of its staged cell: the first window is already in
// the tile (stage_seg_fill with skip 0 returned `kept`), later ones are staged again from the ranges,
// box and visit count stage_seg_ranges left in `aux`. For each window of this segment, on(w, nw, trips, hits, a) gets the window's entry count, this
// lane's trip count, acceptance bits (bit trips - 1 - m for trip m) and sum of accepted terms. Returns
// whether the kept list fit in kMaxWindows windows (otherwise nothing was called). Called by the whole
// wave; leaves the tile free for the next staging.
template <int W, class On>
__device__ inline bool sweep_windows(const Lanes<W>& L, uint32_t kept, float h2_keep, const float4* __restrict__ spos,
                                     uint32_t depth, float4* tile, const StageAux* aux, const float4& pi, uint32_t self,
                                     float h2, float cut, On&& on)
{
    constexpr uint32_t P = W / 8;
    const bool fits = kept != kStageOverflow && kept <= kMaxWindows * depth;
    const uint32_t nwin = fits ? (kept + depth - 1) / depth : 0u;
    const uint32_t wwin = seg_wave_max<W>(nwin);
    for (uint32_t w = 0; w < wwin; w++) {
        if (w > 0) {  // the cell's box and visits from the first staging
            const float3 lo = make_float3(aux->box[0], aux->box[1], aux->box[2]);
            const float3 hi = make_float3(aux->box[3], aux->box[4], aux->box[5]);
            stage_seg_fill<W>(L, w < nwin ? aux->total : 0u, lo, hi, h2_keep, spos, w * depth, depth, tile, aux);
        }
        const uint32_t nw = w < nwin ? min(kept - w * depth, depth) : 0u;
        const uint32_t wtrips = wave_trips<W>(nw);
        const uint32_t trips = (nw + P - 1) / P;
        pad_tile<W>(L, tile, nw, wtrips * P);
        wave_sync();
        float a = 0.f;
        uint64_t hits = accept_sweep<P>(tile + L.ls % P, wtrips, pi, self, h2, cut, a);
        // entries past a segment's end were end markers (never accepted): drop their bits
        hits >>= min(wtrips - trips, 63u);  // trips = 0 leaves no bit set
        if (w < nwin) on(w, nw, trips, hits, a);
        wave_sync();  // the next window overwrites the tile
    }
    return fits;
}

}  // namespace sph
}  // namespace mfl
