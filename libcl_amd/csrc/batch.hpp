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

    // launch shape (mfl_tuning): LIST (group traversal in sorted order) or TILE_LIST (the single-scene
    // tile density pass in a per-scene Morton order); recorded-visit forces either way
    int neighbour_pass = MFL_PASS_TILE_KEPT;
    int density_lanes = 4;
    uint32_t tile_depth = 192;
    uint32_t block_waves = 2;
    uint32_t force_chunks = 2;
    uint32_t refresh_interval = 16;  // steps between traversal order refreshes
    int get_tuning(mfl_tuning* out) const;
    int set_tuning(const mfl_tuning& t);
    int autotune(hipStream_t s, uint32_t steps, mfl_tuning* best, double* ms_per_step);

private:
    int enqueue_step(hipStream_t s);
    int launch_graph(hipStream_t s);
    void drop_graph();
    int refresh_order(hipStream_t s);
    DeviceBuffer* select(int kind);
    bool porder_valid_ = false;
    uint32_t steps_since_refresh_ = 0;
    DeviceBuffer inv_, porder_, mkey_;
    DeviceBuffer kept_, hitmask_, hitx_, meta_;  // TILE_KEPT: per-segment kept entries, acceptance bits, meta
    bool fit_kept();                      // allocates them on first use

    bool ok_ = false;
    uint32_t S_ = 0, N_ = 0, scene_bits_ = 0;
    std::vector<uint32_t> off_host_;
    std::vector<mfl_params> params_host_;
    std::vector<float> box_host_;  // per scene: min xyzw, max xyzw
    bool params_dirty_ = true;
    DeviceBuffer off_, pscene_, params_, boxes_, cuts_;  // cuts_: sqrt_cut(cellSize) per scene
    DeviceBuffer pos_, vel_, force_, cell_, index_, spos_, svel_, tables_, nbr_, ncount_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    std::string graph_sig_;
};

}  // namespace mfl
