// Stable LSD key/value radix sort for gfx950 (the oclRadixSort program, libCL/sort/oclRadixSort.h).
#pragma once

#include "mfl_common.hpp"

namespace mfl {

// Tile shape of one pass: 256 threads (4 wave64s) x `items` keys per thread.
struct SortTuning {
    int items = 0;        // 0 = pick by size
};

struct SortCountPlan {
    uint32_t* counts;
    uint32_t ntiles, tile_keys, bits;
    int items;  // keys per thread of a tile of 256 threads
};

class RadixSort {
public:
    RadixSort() = default;
    ~RadixSort();
    RadixSort(const RadixSort&) = delete;
    RadixSort& operator=(const RadixSort&) = delete;

    // oclProgram::compile analogue: checks that the current device is a gfx950 HIP device.
    int compile();

    // oclRadixSort::compute (libCL/sort/oclRadixSort.cpp:62-144): stable in-place sort of n pairs by
    // bits [start_bit, end_bit). `identity_vals` treats vals[i] == i on input without reading them
    // (the clHash index, libCL/phys/oclFluid3D.cl:63).
    int compute(uint32_t* keys, uint32_t* vals, size_t n, int start_bit, int end_bit, hipStream_t stream,
                bool identity_vals = false);

    // compute() without any device allocation: the temporaries live in a fixed device array of the
    // library (one per GPU, MFL_SORT_NOALLOC_MAX_PAIRS pairs at most; larger n is rejected). Calls that
    // overlap in time on one GPU share that array and race.
    int compute_noalloc(uint32_t* keys, uint32_t* vals, size_t n, int start_bit, int end_bit, hipStream_t stream);

    // Same sort for any bit range in [0, 32] (no multiple-of-4 rule); used by the batched and
    // oclFluid3Dnext steps and the fluid step's traversal order.
    // `first_counted`: the caller already filled the first pass's count table (first_pass_plan).
    int sort_bits(uint32_t* keys, uint32_t* vals, size_t n, int start_bit, int end_bit, hipStream_t stream,
                  bool identity_vals, bool first_counted = false);

    // The first pass's count table, for a caller that fills it while producing the keys (the batched
    // and oclFluid3Dnext hash kernels): tile t holds keys [t tile_keys, (t + 1) tile_keys) and
    // counts[d ntiles + t] must be the number of its keys whose bits [start_bit, start_bit + bits)
    // equal d. Reserves the sort's temporaries.
    int first_pass_plan(size_t n, int start_bit, int end_bit, SortCountPlan* plan);

    // Allocates every temporary a sort of n keys over `nbits` bits needs (so that a later sort
    // enqueues no allocation, e.g. under hipGraph capture).
    int reserve(size_t n, int nbits);

    // Appends the addresses of the sort's temporaries to a launch signature: a captured graph holds
    // them, so a graph must be captured again once any of them is re-allocated.
    void sig_append(std::string& sig) const;

    SortTuning tuning;

private:
    DeviceBuffer tmp_keys_[2], tmp_vals_[2];
    DeviceBuffer counts_, totals_;
};

}  // namespace mfl
