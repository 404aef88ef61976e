// Reverse-mode derivative of an oclFluid3D rollout (an addition; the reference has no adjoint).
#pragma once

#include <vector>

#include "fluid3d.hpp"

namespace mfl {

class Adjoint {
public:
    Adjoint(Fluid3D& fluid, uint32_t steps, uint32_t checkpoint_every, uint32_t list_cap);
    ~Adjoint();
    Adjoint(const Adjoint&) = delete;
    Adjoint& operator=(const Adjoint&) = delete;

    bool ok() const { return ok_; }
    // Runs `steps` steps of the program from its current state, keeping a checkpoint (positions,
    // velocities, forces, bucket tables) every `checkpoint_every` steps and the final state.
    int forward(hipStream_t s);
    // Given dL/dx_T and dL/dv_T (caller-order float4, device, null = 0), walks the rollout backwards
    // segment by segment (recompute from the checkpoint, then reverse each step) and leaves dL/dx_0,
    // dL/dv_0 and the parameter gradients. The program is left in its final state.
    int backward(const float4* pos_bar, const float4* vel_bar, hipStream_t s);
    int get_buffer(int which, void** ptr, size_t* bytes);
    int param_grads(double* out4);

private:
    int save_checkpoint(uint32_t c, hipStream_t s);
    int restore_checkpoint(uint32_t c, hipStream_t s);
    int record_step(uint32_t slot, hipStream_t s);
    int reverse_step(uint32_t slot, hipStream_t s);

    Fluid3D& f_;
    bool ok_ = false;
    bool have_forward_ = false;
    uint32_t steps_ = 0, every_ = 0, cap_ = 0;
    size_t n_ = 0;
    std::vector<DeviceBuffer> ckpt_;          // steps/every + 1 checkpoints (last = final state)
    DeviceBuffer t_index_, t_spos_, t_svel_, t_ncount_, t_nbr_, t_fpre_, t_clean_;  // tape of one segment
    DeviceBuffer xbar_, vbar_, xb_s_, vb_s_, rp_s_, fbar_s_, accbar_, grads_, flags_;
};

}  // namespace mfl
