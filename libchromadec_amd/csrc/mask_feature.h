// Magnitude feature of a tile bin, the learned soft mask's input: shared by
// the tile kernel (which writes it for the paired inference) and the network
// kernels, so that both compute it with the same instructions.
#pragma once
#include <hip/hip_runtime.h>

namespace chd {

// (ln(|a| + 1) - 10) / 3, from the hardware square root and base-2
// logarithm (the argument is >= 1: no denormals).  Its error, a few 1e-7, is
// far below the bf16 step the feature is stored in.
__device__ __forceinline__ float mag_feature(float re, float im)
{
    const float m = __builtin_amdgcn_sqrtf(re * re + im * im);
    return fmaf(__builtin_amdgcn_logf(m + 1.0f), 0.23104906f /* ln 2 / 3 */, -10.0f / 3.0f);
}

}  // namespace chd
