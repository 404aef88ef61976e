// The oclFluid3D program (libCL/phys/oclFluid3D.h:21-123) on gfx950.
#pragma once

#include "mfl_common.hpp"
#include "radix_sort.hpp"

#include <vector>

namespace mfl {

// HIP-event timing of the step's stages (oclKernel::profile / getStartTime / getEndTime,
// libCL/oclKernel.cpp:84-102,167-203).
class StageTimer {
public:
    ~StageTimer();
    uint32_t mask = 0;
    bool begin(int stage, hipStream_t s);   // records a start event when `stage` is enabled
    void end(int stage, hipStream_t s);
    int read(int stage, double* ms, uint64_t* launches, bool reset);

private:
    struct Pending { int stage; hipEvent_t a, b; };
    hipEvent_t take();
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    hipEvent_t open_[MFL_STAGE_COUNT] = {};
    double ms_[MFL_STAGE_COUNT] = {};
    uint64_t count_[MFL_STAGE_COUNT] = {};
    bool drain();
};

class Fluid3D {
public:
    Fluid3D();
    ~Fluid3D();
    Fluid3D(const Fluid3D&) = delete;
    Fluid3D& operator=(const Fluid3D&) = delete;

    int compile();                                  // oclFluid3D::compile, cpp:273-285
    int compute(hipStream_t s);                     // oclFluid3D::compute, cpp:293-342
    int set_particle_count(size_t n);               // cpp:92-124
    size_t particle_count() const { return n_; }    // cpp:126-129
    int set_buffer(int kind, void* ptr, size_t bytes);  // cpp:131-165
    int get_buffer(int kind, void** ptr, size_t* bytes);
    int write_buffer(int kind, const void* host, size_t bytes);
    int read_buffer(int kind, void* host, size_t bytes);
    int enqueue_kernel(const char* name, hipStream_t s);  // getKernel(name) + enqueue

    mfl_params params{};
    float box_min[4] = {-1.f, -1.f, -1.f, 0.f};     // bindBuffers, cpp:260
    float box_max[4] = {1.f, 1.f, 1.f, 0.f};        // cpp:261
    float gravity[4] = {0.f, 0.f, 1.f, 0.f};        // cpp:256
    uint32_t clear_count = kReferenceClear;          // cpp:309-313
    mfl_integrate_fn integrate_cb = nullptr;
    void* integrate_user = nullptr;
    mfl_fluid3d* handle = nullptr;                   // C handle passed back to the callback
    bool use_graph = false;
    bool record_counts = false;             // density passes write each particle's accepted visits
    int traversal = MFL_TRAVERSAL_MORTON;   // order in which the neighbour passes visit particles
    uint32_t refresh_interval = 16;         // steps between Morton order refreshes
    bool traversal_chosen = false;          // set by the caller: set_particle_count keeps it
    // launch shape (mfl_tuning); the defaults are the fastest shape measured on the 1M-particle step
    int neighbour_pass = MFL_PASS_TILE_KEPT;  // LIST, TILE, TILE_LIST or TILE_KEPT neighbour passes
    int density_lanes = 2;                  // lanes per particle in the density pass
    uint32_t tile_depth = 96;               // tile passes: LDS entries per staged cell
    uint32_t block_waves = 2;               // tile passes: waves per workgroup
    uint32_t force_chunks = 2;              // list force pass: 16-byte list chunks per trip
    RadixSort sort;
    StageTimer timer;

    // hash, sort, reorder + bucket clear and bounds of one step (the part before the density pass)
    int enqueue_front(hipStream_t s);
    int tile_items(size_t n) const;
    DeviceBuffer* select(int kind);
    bool check_user_buffers();
    int get_tuning(mfl_tuning* out) const;
    int set_tuning(const mfl_tuning& t);
    int autotune(hipStream_t s, uint32_t steps, mfl_tuning* best, double* ms_per_step);
    // the Morton order is rebuilt by the next compute
    void invalidate_order()
    {
        porder_valid_ = false;
        steps_since_refresh_ = 0;
    }

private:
    bool allocate_internal();
    int enqueue_step(hipStream_t s);
    int enqueue_tile_passes(hipStream_t s, int order);
    int enqueue_list_forces(hipStream_t s, int order);
    int launch_graph(hipStream_t s);
    void drop_graph();

    size_t n_ = kGroup;                              // mParticleCount default, cpp:49
    DeviceBuffer pos_, vel_, force_;
    DeviceBuffer cell_, index_, spos_, svel_, cell_start_, cell_end_, params_buf_;
    DeviceBuffer nbr_, ncount_;  // per-step neighbour visits recorded by the density pass
    DeviceBuffer kept_, hitmask_, hitx_, meta_;  // per-step kept entries and acceptance bits (TILE_KEPT)
    DeviceBuffer inv_;           // inverse of the sort permutation: inv[index[i]] = i
    DeviceBuffer counts_;        // accepted visits per sorted particle (written when record_counts)
    DeviceBuffer porder_, mkey_; // Morton traversal order (caller indices) and its sort keys
    DeviceBuffer skey_;          // small steps: the cell keys in caller order, for the selecting run pass
    DeviceBuffer tkey_, tval_;   // larger steps: keys and particle indices, each hash tile ordered by high digit
    DeviceBuffer tcount_, tstart_;  // per high digit and tile: the tile's count of that digit, its start
    bool porder_valid_ = false;
    uint32_t steps_since_refresh_ = 0;
    uint32_t morton_bits() const;
    int refresh_order(hipStream_t s);

    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    std::string graph_sig_;
};

}  // namespace mfl
