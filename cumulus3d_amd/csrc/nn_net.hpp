// nn_net.hpp — what the PatchStereo session needs of the learned initialiser (nn_net.cpp): its
// sizes, the device feature slots and the forward pass on the backend stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/c3d_init.h"

namespace c3d {
namespace nn {

struct NetView {
    c3d_pm_backend* be;
    int B, H, W;
};

NetView net_view(const c3d_init_net* net);
// device pointer of the [H][W][64] features of `slot`
uint16_t* net_features(c3d_init_net* net, int slot);
// depth range of a slot (decoding of the inverse-depth output)
void net_set_range(c3d_init_net* net, int slot, float dmin, float dmax);
// forward pass over slots [0, batch) queued on the backend stream; returns a c3d_status
int net_forward(c3d_init_net* net, int batch);
// head output [slot][H][W][32] (channels 0..4: inverse-depth logit, normal, confidence logit)
const float* net_output(const c3d_init_net* net);

}  // namespace nn
}  // namespace c3d
