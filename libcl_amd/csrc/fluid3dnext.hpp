// The oclFluid3Dnext program (libCL/phys/oclFluid3Dnext.h:20-130) on gfx950: a 2D particle state
// (x, y, vx, vy) on a dense cellCountX x cellCountY grid.
#pragma once

#include "mfl_common.hpp"
#include "radix_sort.hpp"

namespace mfl {

class Fluid3DNext {
public:
    Fluid3DNext();
    Fluid3DNext(const Fluid3DNext&) = delete;
    Fluid3DNext& operator=(const Fluid3DNext&) = delete;

    int compile();                               // oclFluid3Dnext::compile, cpp:231-255
    int compute(hipStream_t s);                  // oclFluid3Dnext::compute, cpp:264-317
    int set_particle_count(size_t n);            // cpp:85-106
    size_t particle_count() const { return n_; }
    int set_buffer(int kind, void* ptr, size_t bytes);  // setPositionBuffer, cpp:113-123
    int get_buffer(int kind, void** ptr, size_t* bytes);
    int write_buffer(int kind, const void* host, size_t bytes);
    int read_buffer(int kind, void* host, size_t bytes);

    mfl_next_params params{};

private:
    DeviceBuffer* select(int kind);
    bool allocate_internal();
    bool fit_tables();

    RadixSort sort_;
    DeviceBuffer state_, sorted_, cell_, index_, start_, end_, pressure_, force_, params_buf_;
    size_t n_ = 16384;        // 128 * 128, the reference's initial count (cpp:53)
    uint32_t ncells_ = 256;   // mCellCount before compile (cpp:56)
    bool compiled_ = false;
    mfl_next_params uploaded_{};  // the Params last copied to params_buf_
    bool params_synced_ = false;
};

}  // namespace mfl
