// Batched scenes: many independent oclFluid3D systems advanced by one set of launches per step.
#pragma once

#include <vector>

#include "mfl_common.hpp"
#include "radix_sort.hpp"

namespace mfl {

class FluidBatch {
public:
    FluidBatch(uint32_t num_scenes, const uint32_t* counts);
    ~FluidBatch();
    FluidBatch(const FluidBatch&) = delete;
    FluidBatch& operator=(const FluidBatch&) = delete;

    bool ok() const { return ok_; }
    int compute(hipStream_t s);
    int get_buffer(int kind, void** ptr, size_t* bytes);
    int write_buffer(int kind, const void* host, size_t bytes);
    int read_buffer(int kind, void* host, size_t bytes);
    int scene_range(uint32_t scene, uint32_t* offset, uint32_t* count) const;
    int set_params(int scene, const mfl_params& p);
    int get_params(uint32_t scene, mfl_params* p) const;
    int set_box(int scene, const float* mn, const float* mx);

    uint32_t num_scenes() const { return S_; }
    uint32_t total() const { return N_; }

    float gravity[4] = {0.f, 0.f, 1.f, 0.f};
    uint32_t clear_count = kReferenceClear;
    bool use_graph = false;
    RadixSort sort;

private:
    int enqueue_step(hipStream_t s);
    int launch_graph(hipStream_t s);
    void drop_graph();
    DeviceBuffer* select(int kind);

    bool ok_ = false;
    uint32_t S_ = 0, N_ = 0, scene_bits_ = 0;
    std::vector<uint32_t> off_host_;
    std::vector<mfl_params> params_host_;
    std::vector<float> box_host_;  // per scene: min xyzw, max xyzw
    bool params_dirty_ = true;
    DeviceBuffer off_, pscene_, params_, boxes_;
    DeviceBuffer pos_, vel_, force_, cell_, index_, spos_, svel_, tables_, nbr_, ncount_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    std::string graph_sig_;
};

}  // namespace mfl
