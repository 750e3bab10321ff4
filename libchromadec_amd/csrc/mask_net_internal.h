// Learned carrier-mirror mask: layouts shared by the HIP kernels
// (mask_net_kernels.hip) and the host wrapper (mask_net.cpp).
//
// The network works on the 768 visits v = (z*32 + y)*3 + (x-2) of the
// keep/discard rule of TransformPal3D::applyFilter
// (src/decoders/palcolour/transform_pal_3d.cpp:249-333): grid Z=8 x Y=32 x
// X=3, circular in z and y, zero-padded in x.  Three 3x3x3 convolutions,
// 4 -> 32 (ReLU) -> 32 (ReLU) -> 1 (sigmoid).  Tap t = (dz+1)*9 + (dy+1)*3 +
// (dx+1); weights W[t][cin][cout], flattened as
//   W1 [27][4][32], b1 [32], W2 [27][32][32], b2 [32], W3 [27][32], b3.
//
// Device buffers (per tile, all HBM):
//   spectrum  [5][32][8] float2  (bin x = k+2, ky, kz), the tile kernel's output
//   features  [768][4]  bf16
//   act1/act2 [768][32] bf16     (post-ReLU)
//   logit, mask, w, num, dlogit [768] float
//   dpre1/dpre2 [768][32] bf16   (gradient at the pre-activations)
// Weights for the matrix cores are held as B-operand fragments of
// v_mfma_f32_32x32x16_bf16: fragment j, lane l holds the 8 bf16
// B[16j + 8(l/32) + e][l % 32], e = 0..7, as one uint4.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime_api.h>

namespace mnet {

constexpr int kP = 768;          // visits per tile
constexpr int kCh = 32;          // hidden channels
constexpr int kSpecC = 1280;     // complex bins per tile spectrum (5 x 32 x 8)

constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + 27 * 4 * kCh;
constexpr int kOffW2 = kOffB1 + kCh;
constexpr int kOffB2 = kOffW2 + 27 * kCh * kCh;
constexpr int kOffW3 = kOffB2 + kCh;
constexpr int kOffB3 = kOffW3 + 27 * kCh;
constexpr int kNParams = kOffB3 + 1;   // 32033

constexpr int kK1 = 7;    // k-steps of layer 1 (K = 27*4 = 108, padded to 112)
constexpr int kK2 = 54;   // layer 2 (K = 27*32)
constexpr int kK3 = 2;    // layer 3 (K = 32 channels, or 27 taps padded to 32)
constexpr int kWgradSlots = 64;   // workgroups of the weight-gradient kernels

struct Frags {
    uint4 *w1;    // [kK1][64]: B[k = t*4 + ci][n = cout]
    uint4 *w2;    // [kK2][64]: B[k = t*32 + ci][n = cout]
    uint4 *w2t;   // [kK2][64]: B[k = t*32 + cout][n = ci]  (transposed, for the data gradient)
    uint4 *w3t;   // [kK3][64]: B[k = ci][n = t]
    uint4 *w3;    // [kK3][64]: B[k = t][n = ci]
};

struct Acts {
    uint16_t *feat, *act1, *act2, *dpre1, *dpre2;
    float *logit, *w, *num, *dlogit, *loss_tile;
};

void launch_arrange(const float *theta, const Frags &f, hipStream_t s);
void launch_features(const float2 *spec, int64_t n_tiles, uint16_t *feat, hipStream_t s);
void launch_targets(const float2 *spec, const float2 *truth, int64_t n_tiles, float *w, float *num, hipStream_t s);
void launch_forward(const float *theta, const Frags &f, const Acts &a, int64_t n_tiles, float *mask, hipStream_t s);
// Fused, persistent inference (features through sigmoid) for decoding: two
// workgroups per CU (68 KB of LDS each) loop over tiles.
// in: complex tile spectra ([tile][1280] float2).  paired: every tile
// alternates in line parity, so each work unit infers the halves z < 4 of
// two tiles (a tile's mask repeats at (z + 4, y + 16)); `in` is then the
// magnitude features of rows ky < 16 ([tile][640] floats, the tile kernel's
// kTileFeatures mode), and the masks are those of the visits of planes z < 4
// ([tile][384]: visit (z + 4, y + 16) has the mask of visit (z, y))
void launch_infer(const void *in, int64_t n_tiles, const float *theta, const Frags &f, float *mask, int cus,
                  hipStream_t s, bool paired = false);
void launch_backward(const float *theta, const Frags &f, const Acts &a, const float *mask, int64_t n_tiles,
                     float scale, float *partial, float *grad, hipStream_t s);
void launch_adam(float *theta, const float *grad, float *m, float *v, float lr, float beta1, float beta2, float eps,
                 int step, hipStream_t s);

}  // namespace mnet
