// Stable ranking of a tile's keys by one digit with wave64 ballot matching, shared by the radix sort's
// scatter and the fluid step's hash kernel. Keys of wave w, round r, lane l sit at tile offset
// (w ITEMS + r) 64 + l, so (wave, round, lane) order is input order; a key's rank in the tile is
//   [same digit in earlier waves] + [same digit in earlier rounds of its wave] + [lower lanes, same round].
#pragma once

#include "mfl_common.hpp"

namespace mfl {

// rank[r] = the key's rank among its wave's keys of the same digit (in input order); s_cnt[wave * R + d]
// (zeroed beforehand) ends as the wave's count of digit d.
template <int ITEMS, int BITS, uint32_t R, int MAX_BITS>
__device__ __forceinline__ void rank_tile(const uint32_t* key, uint32_t n, size_t tile_base, uint32_t first, int shift,
                                 int bits, uint32_t* s_cnt, uint32_t* rank)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const uint32_t mask = (1u << bits) - 1u;
    const uint64_t lower_lanes = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const size_t idx = tile_base + first + (size_t)r * kWave;
        const bool valid = idx < n;
        const uint32_t d = (key[r] >> shift) & mask;
        // lanes whose digit differs from this lane's in some bit (or that hold no key): per bit b, the
        // ballot of bit b XOR this lane's bit b replicated over 64 bits
        const uint64_t live = __ballot(valid);
        uint32_t differ_lo = ~(uint32_t)live, differ_hi = ~(uint32_t)(live >> 32);
#pragma unroll
        for (int b = 0; b < (BITS > 0 ? BITS : MAX_BITS); b++) {
            if (BITS > 0 || b < bits) {
                const uint32_t rep = (uint32_t)((int)(d << (31 - b)) >> 31);  // 0 or all ones
                const uint64_t bal = __ballot(rep != 0u);
                differ_lo |= (uint32_t)bal ^ rep;
                differ_hi |= (uint32_t)(bal >> 32) ^ rep;
            }
        }
        const uint64_t same = ~((uint64_t)differ_hi << 32 | differ_lo);
        const int leader = __ffsll((unsigned long long)same) - 1;
        uint32_t base = 0;
        if (valid && lane == leader) {
            base = s_cnt[wave * R + d];
            s_cnt[wave * R + d] = base + (uint32_t)__popcll(same);
        }
        base = __shfl(base, leader & (kWave - 1), kWave);
        rank[r] = base + (uint32_t)__popcll(same & lower_lanes);
    }
}

// Turns column d of s_cnt (one count per wave) into offsets over the waves; returns the tile's count of
// digit d (0 for d >= radix).
template <uint32_t R, int WAVES>
__device__ __forceinline__ uint32_t wave_offsets(uint32_t d, uint32_t radix, uint32_t* s_cnt)
{
    if (d >= radix) return 0u;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        const uint32_t c = s_cnt[w * R + d];
        s_cnt[w * R + d] = run;
        run += c;
    }
    return run;
}

}  // namespace mfl
