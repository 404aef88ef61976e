// TEST INFRASTRUCTURE ONLY (the CPU checker / baseline for the standalone key/value sort).
// std::stable_sort of (key, value) pairs by bits [start_bit, end_bit) of the key: the library routine
// the reference's stable 4-bit LSD sort (libCL/sort/oclRadixSort.cpp:62-144) must agree with.
#include <algorithm>
#include <cstdint>
#include <utility>
#include <vector>

extern "C" void orc_std_stable_sort(uint32_t* key, uint32_t* val, uint32_t n, int start_bit, int end_bit)
{
    if (n < 2 || start_bit >= end_bit) return;
    const int w = end_bit - start_bit;
    const uint32_t mask = w >= 32 ? 0xFFFFFFFFu : ((1u << w) - 1u);
    std::vector<std::pair<uint32_t, uint32_t>> kv(n);
    for (uint32_t i = 0; i < n; i++) kv[i] = {key[i], val[i]};
    std::stable_sort(kv.begin(), kv.end(), [&](const std::pair<uint32_t, uint32_t>& a,
                                               const std::pair<uint32_t, uint32_t>& b) {
        return ((a.first >> start_bit) & mask) < ((b.first >> start_bit) & mask);
    });
    for (uint32_t i = 0; i < n; i++) {
        key[i] = kv[i].first;
        val[i] = kv[i].second;
    }
}
