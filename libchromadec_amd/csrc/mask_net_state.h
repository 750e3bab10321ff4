// Internal view of a learned-mask network handle, shared by its API
// (mask_net.cpp) and the device decode (tpal_device.cpp).
#pragma once

#include <chromadec/tpal_mask.h>

#include "mask_net_internal.h"

struct chd_mask_net {
    int device = 0;
    int64_t max_tiles = 0;
    float *theta = nullptr;                  // FP32 parameters
    float *adam_m = nullptr, *adam_v = nullptr;
    float *partial = nullptr;                // weight-gradient partials per workgroup
    float *mask = nullptr;                   // [max_tiles][768] for the gradient pass
    int step = 0;                            // Adam steps taken
    int cus = 256;                           // compute units: persistent inference grid
    mnet::Frags f{};
    mnet::Acts a{};
};
