// Learned carrier-mirror mask for the 3D Transform-PAL tile on gfx950.
//
// The hard keep/discard rule of TransformPal3D::applyFilter
// (src/decoders/palcolour/transform_pal_3d.cpp:249-333) compares each bin
// with its mirror about the U/V carrier.  Here a small convolutional network
// looks at both magnitudes of every visit and its neighbours in (z, y, x)
// and predicts a soft mask in [0, 1] instead.  Layers are implicit GEMMs on
// the matrix cores (v_mfma_f32_32x32x16_bf16: BF16 operands, FP32
// accumulation); activations are stored as BF16 between layers.
//
// Operand layouts of v_mfma_f32_32x32x16_bf16 (64 lanes, l = lane):
//   A (32 x 16): lane l holds A[l % 32][8 (l/32) + e], e = 0..7
//   B (16 x 32): lane l holds B[8 (l/32) + e][l % 32]
//   C (32 x 32): lane l, value v holds C[8 (v/4) + 4 (l/32) + v % 4][l % 32]
//
// One workgroup of 4 waves handles one tile (768 visits = 24 row blocks of
// 32, six per wave) in the forward and data-gradient kernels.  The weight
// gradients accumulate over tiles in registers per workgroup and are
// summed over workgroups in a fixed order, so every result is
// deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "mask_feature.h"
#include "mask_net_internal.h"

namespace mnet {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// float -> bf16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ uint32_t pack2(float lo, float hi)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}

typedef short s16x2 __attribute__((ext_vector_type(2)));

// ReLU and BF16 of two values: converted first, then the negative ones
// cleared as signed 16-bit integers (a BF16 with its sign bit set is a
// negative int16), one packed integer max for the pair.  The bits are those
// of converting the ReLU'd values (rounding keeps the sign).
__device__ __forceinline__ uint32_t pack2_relu(float lo, float hi)
{
    const s16x2 v = __builtin_bit_cast(s16x2, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, (s16x2){0, 0}));
}

// Magnitude feature of a bin: (ln(|a| + 1) - 10) / 3 (mask_feature.h).
__device__ __forceinline__ float mag_feature(float2 a) { return chd::mag_feature(a.x, a.y); }

__device__ __forceinline__ f32x16 mfma(uint4 a, uint4 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// Neighbour of visit p = (z*32 + y)*3 + x at tap t, or -1 where x leaves
// the grid (z and y wrap).
__device__ __forceinline__ int nbr(int p, int t)
{
    const int zy = p / 3, x = p - 3 * zy;
    const int y = zy & 31, z = zy >> 5;
    const int tz = t / 9, r = t - 9 * tz, ty = r / 3, tx = r - 3 * ty;
    const int xs = x + tx - 1;
    if (xs < 0 || xs > 2) return -1;
    return ((((z + tz - 1) & 7) << 5) | ((y + ty - 1) & 31)) * 3 + xs;
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the 256 threads of a workgroup in a fixed order.
__device__ __forceinline__ float block_sum(float v, float *red)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const float s = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// Weight fragments from the FP32 parameters.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(64) k_arrange(const float *__restrict__ th, Frags f)
{
    const int j = blockIdx.x, l = threadIdx.x;
    const int n = l & 31, k0 = 8 * (l >> 5);
    float v[8];
    uint4 *dst;
    if (j < kK1) {  // W1: k = t*4 + ci (k < 108), n = cout
        for (int e = 0; e < 8; ++e) {
            const int k = 16 * j + k0 + e;
            v[e] = k < 108 ? th[kOffW1 + k * kCh + n] : 0.0f;
        }
        dst = f.w1 + j * 64;
    } else if (j < kK1 + kK2) {  // W2: k = t*32 + ci
        const int jj = j - kK1;
        for (int e = 0; e < 8; ++e) v[e] = th[kOffW2 + (16 * jj + k0 + e) * kCh + n];
        dst = f.w2 + jj * 64;
    } else if (j < kK1 + 2 * kK2) {  // W2 transposed: k = t*32 + cout, n = ci
        const int jj = j - kK1 - kK2;
        for (int e = 0; e < 8; ++e) {
            const int k = 16 * jj + k0 + e, t = k >> 5, co = k & 31;
            v[e] = th[kOffW2 + (t * kCh + n) * kCh + co];
        }
        dst = f.w2t + jj * 64;
    } else if (j < kK1 + 2 * kK2 + kK3) {  // W3 for the head: k = ci, n = t
        const int jj = j - kK1 - 2 * kK2;
        for (int e = 0; e < 8; ++e) v[e] = n < 27 ? th[kOffW3 + n * kCh + 16 * jj + k0 + e] : 0.0f;
        dst = f.w3t + jj * 64;
    } else {  // W3 for the data gradient: k = t, n = ci
        const int jj = j - kK1 - 2 * kK2 - kK3;
        for (int e = 0; e < 8; ++e) {
            const int k = 16 * jj + k0 + e;
            v[e] = k < 27 ? th[kOffW3 + k * kCh + n] : 0.0f;
        }
        dst = f.w3 + jj * 64;
    }
    dst[l] = make_uint4(pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7]));
}

// ---------------------------------------------------------------------------
// Features and training targets from tile spectra.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void visit_bins(int p, int &ia, int &ib)
{
    const int zy = p / 3, x = p - 3 * zy;
    const int y = zy & 31, z = zy >> 5;
    const int zr = (10 - z) & 7, yr = (40 - y) & 31;
    ia = (x * 32 + y) * 8 + z;
    ib = ((4 - x) * 32 + yr) * 8 + zr;
}

__global__ void __launch_bounds__(256) k_features(const float2 *__restrict__ spec, uint16_t *__restrict__ feat)
{
    __shared__ float red[4];
    const int64_t t = blockIdx.x;
    const float2 *S = spec + t * kSpecC;
    float f0[3], f1[3], s = 0.0f;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int p = threadIdx.x + 256 * j;
        int ia, ib;
        visit_bins(p, ia, ib);
        const float2 a = S[ia], b = S[ib];
        f0[j] = mag_feature(a);
        f1[j] = mag_feature(b);
        s += f0[j];
    }
    const float mean = block_sum(s, red) * (1.0f / kP);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int p = threadIdx.x + 256 * j;
        // self-mirror visits: x = 4 (index 2), y in {4, 20}, z in {1, 5}
        const int zy = p / 3, x = p - 3 * zy, y = zy & 31, z = zy >> 5;
        const float sm = (x == 2 && (y == 4 || y == 20) && (z == 1 || z == 5)) ? 1.0f : 0.0f;
        reinterpret_cast<uint2 *>(feat)[t * kP + p] = make_uint2(pack2(f0[j], f1[j]), pack2(mean, sm));
    }
}

__global__ void __launch_bounds__(256) k_targets(const float2 *__restrict__ spec, const float2 *__restrict__ truth,
                                                 float *__restrict__ w, float *__restrict__ num)
{
    const int64_t t = blockIdx.x;
    const float2 *S = spec + t * kSpecC, *C = truth + t * kSpecC;
    for (int p = threadIdx.x; p < kP; p += 256) {
        int ia, ib;
        visit_bins(p, ia, ib);
        const float2 a = S[ia], ca = C[ia];
        float ww = a.x * a.x + a.y * a.y;
        float nn = ca.x * a.x + ca.y * a.y;
        if (p % 3 < 2) {  // x = 2, 3: the visit also owns its partner bin
            const float2 b = S[ib], cb = C[ib];
            ww += b.x * b.x + b.y * b.y;
            nn += cb.x * b.x + cb.y * b.y;
        }
        w[t * kP + p] = ww;
        num[t * kP + p] = nn;
    }
}

// ---------------------------------------------------------------------------
// Convolution layers as implicit GEMMs: out[p][n] = sum_k A[p][k] B[k][n].
// ---------------------------------------------------------------------------

enum { kFwd1 = 0, kFwd2 = 1, kBwd2 = 2, kBwd3 = 3 };

// A fragment for row (visit) p of k-step j; hi = lane / 32.
template <int MODE>
__device__ __forceinline__ uint4 a_frag(const void *__restrict__ in, int64_t tbase, int p, int j, int hi)
{
    if constexpr (MODE == kFwd1) {
        // k = t*4 + ci: this lane's 8 k values are taps 4j+2hi and 4j+2hi+1
        const uint2 *f = reinterpret_cast<const uint2 *>(in) + tbase;
        const int t0 = 4 * j + 2 * hi;
        uint2 lo = make_uint2(0u, 0u), up = make_uint2(0u, 0u);
        if (t0 < 27) {
            const int q = nbr(p, t0);
            if (q >= 0) lo = f[q];
        }
        if (t0 + 1 < 27) {
            const int q = nbr(p, t0 + 1);
            if (q >= 0) up = f[q];
        }
        return make_uint4(lo.x, lo.y, up.x, up.y);
    } else if constexpr (MODE == kFwd2 || MODE == kBwd2) {
        // k = t*32 + c; the data gradient reads the mirrored tap 26 - t
        const uint4 *a = reinterpret_cast<const uint4 *>(in) + tbase * 4;  // 4 uint4 per row
        const int t = j >> 1, c0 = (j & 1) * 16 + 8 * hi;
        const int q = nbr(p, MODE == kFwd2 ? t : 26 - t);
        return q >= 0 ? a[q * 4 + (c0 >> 3)] : make_uint4(0u, 0u, 0u, 0u);
    } else {
        // kBwd3: A[q][k = t] = dlogit[nbr(q, -t)], taps 16j + 8hi + e
        const float *d = reinterpret_cast<const float *>(in) + tbase;
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int t = 16 * j + 8 * hi + e;
            const int q = t < 27 ? nbr(p, 26 - t) : -1;
            v[e] = q >= 0 ? d[q] : 0.0f;
        }
        return make_uint4(pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7]));
    }
}

// Forward layers 1 and 2 (bias + ReLU) and the data gradients of layers 3
// and 2 (times the ReLU gate of the layer below).
template <int MODE>
__global__ void __launch_bounds__(256) k_conv(const void *__restrict__ in, const uint4 *__restrict__ wf,
                                              const float *__restrict__ bias, const uint16_t *__restrict__ gate,
                                              uint16_t *__restrict__ out)
{
    constexpr int KS = MODE == kFwd1 ? kK1 : (MODE == kBwd3 ? kK3 : kK2);
    const int64_t t = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 31, hi = lane >> 5;
    const int64_t tbase = t * kP;  // first visit row of the tile
    f32x16 acc[6];
#pragma unroll
    for (int b = 0; b < 6; ++b)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[b][v] = 0.0f;
    for (int j = 0; j < KS; ++j) {
        const uint4 bf = wf[j * 64 + lane];
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            const int p = (wave + 4 * b) * 32 + i;
            acc[b] = mfma(a_frag<MODE>(in, tbase, p, j, hi), bf, acc[b]);
        }
    }
    const float bn = (MODE == kFwd1 || MODE == kFwd2) ? bias[i] : 0.0f;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int p = (wave + 4 * b) * 32 + 8 * (v >> 2) + 4 * hi + (v & 3);
            const int64_t o = (t * kP + p) * kCh + i;
            float x = acc[b][v];
            if constexpr (MODE == kFwd1 || MODE == kFwd2) {
                x = fmaxf(x + bn, 0.0f);
            } else {
                x = bf2f(gate[o]) > 0.0f ? x : 0.0f;
            }
            out[o] = (uint16_t)f2bf(x);
        }
    }
}

// Layer 3 (32 -> 1) and the sigmoid.  G[q][t] = sum_c act2[q][c] W3[t][c]
// on the matrix cores, then logit[p] = b3 + sum_t G[nbr(p, t)][t] in a
// fixed tap order.
__global__ void __launch_bounds__(256) k_head(const uint16_t *__restrict__ act2, const uint4 *__restrict__ w3t,
                                              const float *__restrict__ b3, float *__restrict__ logit,
                                              float *__restrict__ mask)
{
    __shared__ float G[kP * 27];
    const int64_t t = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 31, hi = lane >> 5;
    const uint4 *a = reinterpret_cast<const uint4 *>(act2) + t * kP * 4;
    const uint4 bf0 = w3t[lane], bf1 = w3t[64 + lane];
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const int p = (wave + 4 * b) * 32 + i;
        f32x16 acc;
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
        acc = mfma(a[p * 4 + hi], bf0, acc);
        acc = mfma(a[p * 4 + 2 + hi], bf1, acc);
        if (i < 27) {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int q = (wave + 4 * b) * 32 + 8 * (v >> 2) + 4 * hi + (v & 3);
                G[q * 27 + i] = acc[v];
            }
        }
    }
    __syncthreads();
    const float bias = b3[0];
    for (int p = threadIdx.x; p < kP; p += 256) {
        float s = bias;
        for (int tap = 0; tap < 27; ++tap) {
            const int q = nbr(p, tap);
            if (q >= 0) s += G[q * 27 + tap];
        }
        logit[t * kP + p] = s;
        mask[t * kP + p] = 1.0f / (1.0f + expf(-s));
    }
}

// ---------------------------------------------------------------------------
// Fused inference: features, the three layers and the sigmoid of a tile in
// one workgroup, activations in LDS.  Persistent, two 4-wave workgroups per
// CU on different tiles, so that one's feature, epilogue and head phases
// overlap the other's matrix-core phases.  Layer-2 weight fragments stream
// from the L1/L2 caches (55 KB, read by every workgroup), loaded three
// half-taps ahead.  Same arithmetic as k_features + k_conv<kFwd1/kFwd2> +
// k_head (BF16 operands, FP32 accumulation, BF16 activations), so masks
// agree with the layered forward up to FP32 summation order.
// ---------------------------------------------------------------------------

constexpr int kActStride = 40;  // bf16 per activation row in LDS (32 + 8: spreads gathered rows over banks)
constexpr int kActRowB = 2 * kActStride;
constexpr int kZeroRow = kP;    // an all-zero feature row: the neighbour of a visit across the x edge
constexpr int kHeadSlots = 16;  // head taps per round: taps 0-15, then 16-26
constexpr int kGSlotB = kP * 4; // bytes per head tap slot: G[t][q'] of every visit

struct LdsInfer {
    union {
        uint16_t act[kP * kActStride];         // act1, then act2
        float g[kHeadSlots * kP];              // head partial sums of one round, [tap slot][q']
    } u;
    uint2 feat[kP + 1];                        // 4 bf16 features per visit (+ zero row)
    alignas(16) float bias[2][32];             // layer-1 and layer-2 biases (read as float4)
    uint4 w1[kK1 * 64];                        // layer-1 weight fragments, [k-step][lane]
};

// Row blocks: wave w owns the 6 blocks (z, x) with z in {2w, 2w+1} and x in
// {0, 1, 2}; an MFMA row is y = 0..31, so visit p = (z*32 + y)*3 + x.  A
// neighbour at tap (dz, dy, dx) is then row ((z+dz)&7)*96 + ((y+dy)&31)*3 +
// x+dx: a wave-uniform z term, a per-lane y term and a compile-time x, and
// x+dx outside 0..2 is the zero row, known at compile time.
//
// Paired form (HALF): when the tiles alternate in line parity, a tile's
// features, and so its activations and mask, repeat half a period away in z
// and y: visit (z + 4, y + 16, x) equals visit (z, y, x) (its bins are the
// same magnitudes), and the convolutions, circular in z and y, keep that
// symmetry.  So the grid holds the halves z < 4 of two tiles, the first in
// z rows 0-3 and the second in rows 4-7, and a neighbour that leaves a
// half, z - 1 of z = 0 or z + 1 of z = 3, is the half's row z + 3 or z - 3
// at y + 16.  A wave's four neighbour planes z0 - 1 .. z0 + 2 leave its half
// at one end only: plane 0 for the even waves, plane 3 for the odd ones.
constexpr int kInferWaves = 4;                   // 256 threads; two workgroups per CU
constexpr int kInferBlocks = 24 / kInferWaves;   // (z, x) row blocks per wave
constexpr int kInferVisits = 3;                  // feature/head visits per thread: x = 0..2

// Per-lane y terms of the neighbour planes: mid for planes 1 and 2, lo and hi
// for planes 0 and 3 (the same as mid, or shifted by 16 rows where the plane
// is the other end of a paired half).
struct YTerms {
    int lo[3], mid[3], hi[3];
};
template <int ZI>
__device__ __forceinline__ int yterm(const YTerms &y, int d)
{
    return ZI == 0 ? y.lo[d] : ZI == 3 ? y.hi[d] : y.mid[d];
}

// Byte offset of act row (zt, lane y term, x) for tap T of block B (compile
// time); zt[k] = ((z0 + k - 1) & 7) * 96 * rowbytes for k = 0..3 covers
// z0-1 .. z0+2 (z0 = 2w, blocks z0 and z0+1).
template <int B, int T, int ROWB>
__device__ __forceinline__ int tap_addr(const int (&zt)[4], const YTerms &yt, int zero)
{
    constexpr int dz = T / 9 - 1, dy = (T / 3) % 3 - 1, dx = T % 3 - 1;
    constexpr int x = B % 3 + dx, zi = B / 3 + dz + 1;
    if constexpr (x < 0 || x > 2) return zero;
    else return zt[zi] + yterm<zi>(yt, dy + 1) + x * ROWB;
}

// Layer-1 k-step J: this lane's A operand, the features of its two taps
// 4J + 2hi and 4J + 2hi + 1 (taps >= 27 are K padding: the zero row).
template <int B, int T>
__device__ __forceinline__ int l1_addr(const int (&zt)[4], const YTerms &yt, int zero)
{
    if constexpr (T >= 27) return zero;
    else return tap_addr<B, T, 8>(zt, yt, zero);
}

template <int J, int B>
__device__ __forceinline__ uint4 l1_operand(const char *fb, const int (&zt)[4], const YTerms &yt, int zero, int hi)
{
    const int qa = hi ? l1_addr<B, 4 * J + 2>(zt, yt, zero) : l1_addr<B, 4 * J>(zt, yt, zero);
    const int qb = hi ? l1_addr<B, 4 * J + 3>(zt, yt, zero) : l1_addr<B, 4 * J + 1>(zt, yt, zero);
    const uint2 lo = *reinterpret_cast<const uint2 *>(fb + qa), up = *reinterpret_cast<const uint2 *>(fb + qb);
    return make_uint4(lo.x, lo.y, up.x, up.y);
}

// Layer-1 k-step J: the matrix op of each block, then that block's gather
// for step J+1 into the same registers; step J+1's weight fragment (LDS,
// w1l = this lane's word of step 0) is read while step J's ops run.
template <int J>
__device__ __forceinline__ void layer1_steps(const char *fb, const int (&zt)[4], const YTerms &yt, int zero, int hi,
                                             const char *w1l, uint4 &bf, uint4 (&af)[kInferBlocks],
                                             f32x16 (&acc)[kInferBlocks])
{
    if constexpr (J < kK1) {
        constexpr bool more = J + 1 < kK1;
        const uint4 bcur = bf;
        if constexpr (more) bf = *reinterpret_cast<const uint4 *>(w1l + (J + 1) * 1024);
        acc[0] = mfma(bcur, af[0], acc[0]);
        if constexpr (more) af[0] = l1_operand<J + 1, 0>(fb, zt, yt, zero, hi);
        acc[1] = mfma(bcur, af[1], acc[1]);
        if constexpr (more) af[1] = l1_operand<J + 1, 1>(fb, zt, yt, zero, hi);
        acc[2] = mfma(bcur, af[2], acc[2]);
        if constexpr (more) af[2] = l1_operand<J + 1, 2>(fb, zt, yt, zero, hi);
        acc[3] = mfma(bcur, af[3], acc[3]);
        if constexpr (more) af[3] = l1_operand<J + 1, 3>(fb, zt, yt, zero, hi);
        acc[4] = mfma(bcur, af[4], acc[4]);
        if constexpr (more) af[4] = l1_operand<J + 1, 4>(fb, zt, yt, zero, hi);
        acc[5] = mfma(bcur, af[5], acc[5]);
        if constexpr (more) af[5] = l1_operand<J + 1, 5>(fb, zt, yt, zero, hi);
        __builtin_amdgcn_sched_barrier(0);
        layer1_steps<J + 1>(fb, zt, yt, zero, hi, w1l, bf, af, acc);
    }
}

// Layer 2 runs in 18 groups G = ((dz+1)*3 + dy+1)*2 + h.  A group gathers
// the six act rows (z0 + zb + dz, y + dy, x') for zb = 0, 1 and x' = 0..2,
// channels 16h .. 16h+15 (byte addresses include this lane's 8-channel half),
// and serves the three taps dx = -1, 0, 1 of both z blocks: block (zb, x)
// takes row x' = x + dx.  Pairs with x + dx outside 0..2 would read the zero
// row and are skipped, so a group is 6 gathers and 14 matrix ops.
constexpr int kL2Groups = 18;
constexpr int kL2Ring = 9;   // weight fragments in flight: three groups

template <int G, int ZB, int XP>
__device__ __forceinline__ uint4 l2_operand(const char *ab, const int (&zt)[4], const YTerms &yt)
{
    constexpr int dz = G / 6 - 1, dy = (G / 2) % 3 - 1, h = G % 2;
    return *reinterpret_cast<const uint4 *>(ab + zt[ZB + dz + 1] + yterm<ZB + dz + 1>(yt, dy + 1) + XP * kActRowB +
                                            32 * h);
}

// Weight half-tap of group G and dx = D - 1: tap ((dz+1)*3 + dy+1)*3 + D,
// half G % 2 (1 KB per half-tap, tap-major).
template <int G, int D>
constexpr int l2_half_tap()
{
    return ((G / 2) * 3 + D) * 2 + G % 2;
}

template <int G>
__device__ __forceinline__ void l2_fetch_weights(const char *w2, int w2off, uint4 (&wq)[kL2Ring])
{
    wq[(3 * G + 0) % kL2Ring] = *reinterpret_cast<const uint4 *>(w2 + w2off + l2_half_tap<G, 0>() * 1024);
    wq[(3 * G + 1) % kL2Ring] = *reinterpret_cast<const uint4 *>(w2 + w2off + l2_half_tap<G, 1>() * 1024);
    wq[(3 * G + 2) % kL2Ring] = *reinterpret_cast<const uint4 *>(w2 + w2off + l2_half_tap<G, 2>() * 1024);
}

// Group G: its matrix ops, each operand regathered for group G+1 as soon as
// its last op is issued.  The two z blocks alternate so that no accumulator
// is hit by consecutive ops.
template <int G>
__device__ __forceinline__ void layer2_groups(const char *ab, const int (&zt)[4], const YTerms &yt,
                                              const char *w2, int w2off, uint4 (&op)[kInferBlocks],
                                              uint4 (&wq)[kL2Ring], f32x16 (&acc)[kInferBlocks])
{
    if constexpr (G < kL2Groups) {
        if constexpr (G + 2 < kL2Groups) l2_fetch_weights<G + 2>(w2, w2off, wq);
        const uint4 wm = wq[(3 * G + 0) % kL2Ring], w0 = wq[(3 * G + 1) % kL2Ring], wp = wq[(3 * G + 2) % kL2Ring];
        constexpr bool more = G + 1 < kL2Groups;
        // x' = 0: block x = 0 (dx 0), x = 1 (dx -1)
        acc[0] = mfma(w0, op[0], acc[0]);
        acc[3] = mfma(w0, op[3], acc[3]);
        acc[1] = mfma(wm, op[0], acc[1]);
        acc[4] = mfma(wm, op[3], acc[4]);
        if constexpr (more) {
            op[0] = l2_operand<G + 1, 0, 0>(ab, zt, yt);
            op[3] = l2_operand<G + 1, 1, 0>(ab, zt, yt);
        }
        // x' = 1: x = 2 (dx -1), x = 0 (dx +1), x = 1 (dx 0)
        acc[2] = mfma(wm, op[1], acc[2]);
        acc[5] = mfma(wm, op[4], acc[5]);
        acc[0] = mfma(wp, op[1], acc[0]);
        acc[3] = mfma(wp, op[4], acc[3]);
        acc[1] = mfma(w0, op[1], acc[1]);
        acc[4] = mfma(w0, op[4], acc[4]);
        if constexpr (more) {
            op[1] = l2_operand<G + 1, 0, 1>(ab, zt, yt);
            op[4] = l2_operand<G + 1, 1, 1>(ab, zt, yt);
        }
        // x' = 2: x = 2 (dx 0), x = 1 (dx +1)
        acc[2] = mfma(w0, op[2], acc[2]);
        acc[5] = mfma(w0, op[5], acc[5]);
        acc[1] = mfma(wp, op[2], acc[1]);
        acc[4] = mfma(wp, op[5], acc[4]);
        if constexpr (more) {
            op[2] = l2_operand<G + 1, 0, 2>(ab, zt, yt);
            op[5] = l2_operand<G + 1, 1, 2>(ab, zt, yt);
        }
        // keep each group's loads next to its matrix ops: hoisting later
        // operands would exceed the register budget
        __builtin_amdgcn_sched_barrier(0);
        layer2_groups<G + 1>(ab, zt, yt, w2, w2off, op, wq, acc);
    }
}

// Partial logit of visit (z, y, X) over head taps T .. E-1 of the round whose
// first tap is R0: G holds slot t - R0 of every visit q' = ((z*3 + x)*32 + y).
// rb[dz+1][dy+1] = ((z+dz)&7)*384 + ((y+dy)&31)*4; taps whose x + dx leaves
// 0..2 read the zero padding of the reference's filter: they are skipped.
template <int X, int T, int E, int R0>
__device__ __forceinline__ float head_sum(const char *g, const int (&rb)[3][3], float s)
{
    if constexpr (T < E) {
        constexpr int dz = T / 9 - 1, dy = (T / 3) % 3 - 1, dx = T % 3 - 1, x = X + dx;
        if constexpr (x >= 0 && x <= 2)
            s += *reinterpret_cast<const float *>(g + rb[dz + 1][dy + 1] + (T - R0) * kGSlotB + x * 128);
        return head_sum<X, T + 1, E, R0>(g, rb, s);
    } else {
        return s;
    }
}

// Accumulators of a transposed matrix op start at the bias: lane (y, hi)
// holds output channels 8j + 4hi + k in acc[b][4j + k].
__device__ __forceinline__ void acc_from_bias(const float *bias, int hi, f32x16 (&acc)[kInferBlocks])
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float4 bv = *reinterpret_cast<const float4 *>(bias + 8 * j + 4 * hi);
#pragma unroll
        for (int b = 0; b < kInferBlocks; ++b) {
            acc[b][4 * j] = bv.x;
            acc[b][4 * j + 1] = bv.y;
            acc[b][4 * j + 2] = bv.z;
            acc[b][4 * j + 3] = bv.w;
        }
    }
}

// ReLU, BF16, and the activation rows: channels 8j + 4hi .. +3 of visit
// (z0 + b/3, y, b%3) are one 8-byte store (ebase: this lane's row of block
// 0 plus its channel half).
__device__ __forceinline__ void store_act(char *act, int ebase, const f32x16 (&acc)[kInferBlocks])
{
#pragma unroll
    for (int b = 0; b < kInferBlocks; ++b)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint2 pk = make_uint2(pack2_relu(acc[b][4 * j], acc[b][4 * j + 1]),
                                        pack2_relu(acc[b][4 * j + 2], acc[b][4 * j + 3]));
            *reinterpret_cast<uint2 *>(act + ebase + ((b / 3) * 96 + b % 3) * kActRowB + 16 * j) = pk;
        }
}

// HALF: the paired form (every tile alternates in line parity); a work unit
// is tiles 2u and 2u + 1 (halves z < 4 in z rows 0-3 and 4-7), and the masks
// of the visits of planes z < 4 are written, [tile][384] (visit (z + 4,
// y + 16) has the mask of visit (z, y)).
// Its input is the magnitude feature of each bin of rows ky < 16
// ([tile][k][ky < 16][kz] floats, written by the tile kernel): a bin of row
// ky >= 16 is read at its image (ky - 16, kz + 4), which has the same
// magnitude.  Otherwise the input is complex tile spectra ([tile][1280]).
template <bool HALF>
__global__ void __launch_bounds__(64 * kInferWaves, 2)
k_infer(const void *__restrict__ in, int64_t n_tiles, const float *__restrict__ theta, Frags f,
        float *__restrict__ mask)
{
    constexpr int NT = 64 * kInferWaves, NB = kInferBlocks, NV = kInferVisits;
    const int64_t n_units = HALF ? (n_tiles + 1) / 2 : n_tiles;
    __shared__ LdsInfer L;
    __shared__ float red[kInferWaves];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int i = lane & 31, hi = lane >> 5;
    const int z0 = 2 * wave;
    if (tid == 0) L.feat[kZeroRow] = make_uint2(0u, 0u);
    if (tid < 64) L.bias[tid >> 5][tid & 31] = theta[(tid < 32 ? kOffB1 : kOffB2) + (tid & 31)];
    for (int k = tid; k < kK1 * 64; k += 64 * kInferWaves) L.w1[k] = f.w1[k];
    const float b3 = theta[kOffB3];
    // address terms (bytes): act rows for layer 2, feature rows for layer 1;
    // neighbour plane k is z row z0 + k - 1, within the wave's half when
    // paired
    int zt_act[4], zt_feat[4], yt_act[3], yt_feat[3], ys_act[3], ys_feat[3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int zr = (HALF ? ((z0 & 4) | ((z0 + k - 1) & 3)) : ((z0 + k - 1) & 7)) * 96;
        zt_act[k] = zr * kActRowB;
        zt_feat[k] = zr * 8;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int yr = ((i + k - 1) & 31) * 3, ys = ((i + k + 15) & 31) * 3;   // y + dy, y + dy + 16
        yt_act[k] = yr * kActRowB + 16 * hi;
        yt_feat[k] = yr * 8;
        ys_act[k] = ys * kActRowB + 16 * hi;
        ys_feat[k] = ys * 8;
    }
    const int zero_feat = kZeroRow * 8;
    // this thread's feature/head visits: z row zv = z0 + hi, y = i, x = 0..2;
    // paired, z row zv is z = zv % 4 of tile 2u + zv / 4
    const int zv = z0 + hi;
    const int pv = (zv * 32 + i) * 3;
    const int zt_true = HALF ? (zv & 3) : zv;
    const int pvt = (zt_true * 32 + i) * 3;   // the visit in its own tile
    int ia[NV], ib[NV];   // byte offsets of the visit's bin and its mirror in a tile spectrum
    auto compact = [](int b) {
        const int k = b >> 8, ky = (b >> 3) & 31, kz = b & 7;
        return ky < 16 ? (k * 16 + ky) * 8 + kz : (k * 16 + ky - 16) * 8 + ((kz + 4) & 7);
    };
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        visit_bins(pvt + j, ia[j], ib[j]);
        if constexpr (HALF) {
            ia[j] = compact(ia[j]);
            ib[j] = compact(ib[j]);
        }
        ia[j] *= HALF ? (int)sizeof(float) : (int)sizeof(float2);
        ib[j] *= HALF ? (int)sizeof(float) : (int)sizeof(float2);
    }
    // the wave's tile within a work unit (paired: waves 0 and 1 hold the
    // first tile's half, waves 2 and 3 the second's)
    const int wtile = HALF ? (wave >> 1) : 0;
    const char *ab = reinterpret_cast<const char *>(L.u.act);
    const char *fb = reinterpret_cast<const char *>(L.feat);
    // the bins (features, paired) of this thread's visits, loaded one tile
    // ahead
    using In = typename std::conditional<HALF, float, float2>::type;
    constexpr int64_t kInBytes = HALF ? (int64_t)(kSpecC / 2) * sizeof(float) : (int64_t)kSpecC * sizeof(float2);
    In sa[NV], sb[NV];
    auto fetch = [&](int64_t u) {
        const int64_t t = HALF ? 2 * u + wtile : u;
        const char *st = reinterpret_cast<const char *>(in) + t * kInBytes;   // wave-uniform
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            sa[j] = sb[j] = In{};
            if (t < n_tiles) {
                sa[j] = *reinterpret_cast<const In *>(st + ia[j]);
                sb[j] = *reinterpret_cast<const In *>(st + ib[j]);
            }
        }
    };
    fetch(blockIdx.x);
    __syncthreads();

    for (int64_t u = blockIdx.x; u < n_units; u += gridDim.x) {
        const int64_t t = HALF ? 2 * u + wtile : u;   // this wave's tile
        // Keep the address terms opaque per tile: otherwise every tap's
        // address (hundreds per lane) is hoisted out of the tile loop.
#pragma unroll
        for (int k = 0; k < 3; ++k) asm volatile("" : "+v"(yt_act[k]), "+v"(yt_feat[k]));
        YTerms ya, yf;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ya.mid[k] = ya.lo[k] = ya.hi[k] = yt_act[k];
            yf.mid[k] = yf.lo[k] = yf.hi[k] = yt_feat[k];
        }
        if constexpr (HALF) {
            // plane 0 (even waves) or plane 3 (odd waves) is the far end of
            // the half, 16 rows on
#pragma unroll
            for (int k = 0; k < 3; ++k) asm volatile("" : "+v"(ys_act[k]), "+v"(ys_feat[k]));
            const bool odd = wave & 1;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                ya.lo[k] = odd ? yt_act[k] : ys_act[k];
                ya.hi[k] = odd ? ys_act[k] : yt_act[k];
                yf.lo[k] = odd ? yt_feat[k] : ys_feat[k];
                yf.hi[k] = odd ? ys_feat[k] : yt_feat[k];
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) asm volatile("" : "+s"(zt_act[k]), "+s"(zt_feat[k]));
#pragma unroll
        for (int j = 0; j < NV; ++j) asm volatile("" : "+v"(ia[j]), "+v"(ib[j]));
        // per-lane byte bases of the epilogue stores and head loads; each
        // store adds a compile-time offset (block, value) to them
        int ebase = ((z0 * 32 + i) * 3) * kActRowB + 8 * hi;           // act row of block 0, channel half
        int gbase = 4 * hi * kGSlotB + (z0 * 96 + i) * 4;                // G slot 4hi, visit q' of block 0
        int hbase = ((z0 * 32 + i) * 3) * kActRowB + 16 * hi;            // head A row of block 0
        int w2off = lane * 16;                                           // weight fragment word
        asm volatile("" : "+v"(ebase), "+v"(gbase), "+v"(hbase), "+v"(w2off));
        // ---- features (as k_features) ----------------------------------
        float f0[NV], f1[NV], s = 0.0f;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            if constexpr (HALF) {
                f0[j] = sa[j];
                f1[j] = sb[j];
            } else {
                f0[j] = mag_feature(sa[j]);
                f1[j] = mag_feature(sb[j]);
            }
            s += f0[j];
        }
        s = wave_sum(s);
        if (lane == 0) red[wave] = s;
        __syncthreads();
        float tot = 0.0f;
        float mean;
        if constexpr (HALF) {
            // the tile's mean over its half (the other half repeats it)
            tot = red[2 * wtile] + red[2 * wtile + 1];
            mean = tot * (2.0f / kP);
        } else {
#pragma unroll
            for (int w = 0; w < kInferWaves; ++w) tot += red[w];
            mean = tot * (1.0f / kP);
        }
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            // self-mirror bins: x = 2 (j), y in {4, 20}, z in {1, 5}
            const float sm = (j == 2 && (i == 4 || i == 20) && (zv == 1 || zv == 5)) ? 1.0f : 0.0f;
            L.feat[pv + j] = make_uint2(pack2(f0[j], f1[j]), pack2(mean, sm));
        }
        __syncthreads();

        // The matrix ops are transposed (weights as the A operand): lane
        // (y, hi) of block b gets output channels 8j + 4hi + k of visit
        // (z0 + b/3, y, b%3) in acc[b][4j + k].
        f32x16 acc[NB];
        // ---- layer 1: K = t*4 + ci (112) --------------------------------
        acc_from_bias(L.bias[0], hi, acc);
        {
            const char *w1l = reinterpret_cast<const char *>(L.w1) + 16 * lane;
            uint4 bf = *reinterpret_cast<const uint4 *>(w1l);
            uint4 af[NB];
            af[0] = l1_operand<0, 0>(fb, zt_feat, yf, zero_feat, hi);
            af[1] = l1_operand<0, 1>(fb, zt_feat, yf, zero_feat, hi);
            af[2] = l1_operand<0, 2>(fb, zt_feat, yf, zero_feat, hi);
            af[3] = l1_operand<0, 3>(fb, zt_feat, yf, zero_feat, hi);
            af[4] = l1_operand<0, 4>(fb, zt_feat, yf, zero_feat, hi);
            af[5] = l1_operand<0, 5>(fb, zt_feat, yf, zero_feat, hi);
            layer1_steps<0>(fb, zt_feat, yf, zero_feat, hi, w1l, bf, af, acc);
        }
        store_act(reinterpret_cast<char *>(L.u.act), ebase, acc);
        // the first two groups' layer-2 weight fragments, in flight across
        // the barrier
        const char *w2 = reinterpret_cast<const char *>(f.w2);
        uint4 wq[kL2Ring];
        l2_fetch_weights<0>(w2, w2off, wq);
        l2_fetch_weights<1>(w2, w2off, wq);
        __syncthreads();

        // ---- layer 2: 27 taps x 32 channels, in groups of 3 taps x 16 ----
        acc_from_bias(L.bias[1], hi, acc);
        {
            uint4 op[NB];
            op[0] = l2_operand<0, 0, 0>(ab, zt_act, ya);
            op[1] = l2_operand<0, 0, 1>(ab, zt_act, ya);
            op[2] = l2_operand<0, 0, 2>(ab, zt_act, ya);
            op[3] = l2_operand<0, 1, 0>(ab, zt_act, ya);
            op[4] = l2_operand<0, 1, 1>(ab, zt_act, ya);
            op[5] = l2_operand<0, 1, 2>(ab, zt_act, ya);
            layer2_groups<0>(ab, zt_act, ya, w2, w2off, op, wq, acc);
        }
        // head weights before the next tile's bins: loads complete in order,
        // so the head then waits only for its own (w2off is opaque per tile,
        // which keeps these loads in the loop)
        const char *w3 = reinterpret_cast<const char *>(f.w3t);
        const uint4 w3a = *reinterpret_cast<const uint4 *>(w3 + w2off);
        const uint4 w3b = *reinterpret_cast<const uint4 *>(w3 + w2off + 1024);
        fetch(u + gridDim.x);   // next tile's bins, in flight through the head
        __syncthreads();  // every wave has read act1
        store_act(reinterpret_cast<char *>(L.u.act), ebase, acc);
        __syncthreads();

        // ---- head: G[q][t] = sum_c act2[q][c] W3[t][c] ------------------
        {
            uint4 a0[NB], a1[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int off = ((b / 3) * 32 * 3 + b % 3) * kActRowB;
                a0[b] = *reinterpret_cast<const uint4 *>(ab + hbase + off);
                a1[b] = *reinterpret_cast<const uint4 *>(ab + hbase + off + 32);
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[b][v] = 0.0f;
                acc[b] = mfma(w3a, a0[b], acc[b]);
                acc[b] = mfma(w3b, a1[b], acc[b]);
            }
        }
        // acc[b][4j + k] is G of tap 8j + 4hi + k (taps 27-31 are zero
        // weights) at visit (z0 + b/3, y, b%3)
        float logit[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) logit[j] = b3;
        const char *gb = reinterpret_cast<const char *>(L.u.g);
        int rb[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            // paired: a neighbour beyond the half is its far end, 16 rows on
            const int zl = (zv & 3) + a - 1;
            const bool wrap = HALF && (zl < 0 || zl > 3);
            const int zn = HALF ? ((zv & 4) | (zl & 3)) : ((zv + a - 1) & 7);
#pragma unroll
            for (int c = 0; c < 3; ++c) rb[a][c] = zn * 384 + ((i + c - 1 + (wrap ? 16 : 0)) & 31) * 4;
        }
        // two rounds through LDS: taps 0-15 (j = 0, 1), then 16-26 (j = 2, 3)
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            __syncthreads();  // act2 (round 0) or the previous round's G is read
            char *gw = reinterpret_cast<char *>(L.u.g) + gbase;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int off = (8 * jj + k) * kGSlotB + ((b / 3) * 3 + b % 3) * 128;
                        *reinterpret_cast<float *>(gw + off) = acc[b][4 * (2 * r + jj) + k];
                    }
            __syncthreads();
            if (r == 0) {
                logit[0] = head_sum<0, 0, 16, 0>(gb, rb, logit[0]);
                logit[1] = head_sum<1, 0, 16, 0>(gb, rb, logit[1]);
                logit[2] = head_sum<2, 0, 16, 0>(gb, rb, logit[2]);
            } else {
                logit[0] = head_sum<0, 16, 27, 16>(gb, rb, logit[0]);
                logit[1] = head_sum<1, 16, 27, 16>(gb, rb, logit[1]);
                logit[2] = head_sum<2, 16, 27, 16>(gb, rb, logit[2]);
            }
        }
        if (t < n_tiles) {
            // paired: the visits of planes z < 4 only ([tile][384]); visit
            // (z + 4, y + 16) has the mask of visit (z, y)
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                const float m = 1.0f / (1.0f + expf(-logit[j]));
                mask[t * (HALF ? kP / 2 : kP) + pvt + j] = m;
            }
        }
        __syncthreads();  // G is read before the next tile's features and act1 overwrite it
    }
}

// ---------------------------------------------------------------------------
// Loss and its gradient at the logits.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_loss(const float *__restrict__ mask, const float *__restrict__ w,
                                              const float *__restrict__ num, float scale,
                                              float *__restrict__ dlogit, float *__restrict__ loss_tile)
{
    __shared__ float red[4];
    const int64_t t = blockIdx.x;
    float l = 0.0f;
    for (int p = threadIdx.x; p < kP; p += 256) {
        const int64_t o = t * kP + p;
        const float m = mask[o], ww = w[o], nn = num[o];
        dlogit[o] = 2.0f * scale * (ww * m - nn) * m * (1.0f - m);
        l += scale * (ww * m * m - 2.0f * nn * m);
    }
    l = block_sum(l, red);
    if (threadIdx.x == 0) loss_tile[t] = l;
}

// ---------------------------------------------------------------------------
// Weight gradients: per workgroup over tiles blockIdx.x, +gridDim.x, ...,
// written to partial[slot][param]; k_reduce sums the slots in order.
// ---------------------------------------------------------------------------

// dW2[t][ci][co] = sum_q act1[q][ci] dpre2[nbr(q, -t)][co]; wave w owns taps
// w, w+4, ... (C rows = ci, columns = co, K = q).
__global__ void __launch_bounds__(256) k_wgrad2(const uint16_t *__restrict__ act1,
                                                const uint16_t *__restrict__ dpre2, int64_t n_tiles,
                                                float *__restrict__ partial)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 31, hi = lane >> 5;
    f32x16 acc[7];
#pragma unroll
    for (int k = 0; k < 7; ++k)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[k][v] = 0.0f;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint16_t *A = act1 + t * kP * kCh;
        const uint16_t *D = dpre2 + t * kP * kCh;
        for (int kb = 0; kb < kP / 16; ++kb) {
            const int q0 = kb * 16 + 8 * hi;
            uint32_t av[4];
#pragma unroll
            for (int e = 0; e < 4; ++e)
                av[e] = (uint32_t)A[(q0 + 2 * e) * kCh + i] | ((uint32_t)A[(q0 + 2 * e + 1) * kCh + i] << 16);
            const uint4 af = make_uint4(av[0], av[1], av[2], av[3]);
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const int tap = wave + 4 * k;
                if (tap < 27) {
                    uint32_t bv[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int q = nbr(q0 + e, 26 - tap);
                        bv[e] = q >= 0 ? (uint32_t)D[q * kCh + i] : 0u;
                    }
                    const uint4 bfr = make_uint4(bv[0] | (bv[1] << 16), bv[2] | (bv[3] << 16),
                                                 bv[4] | (bv[5] << 16), bv[6] | (bv[7] << 16));
                    acc[k] = mfma(af, bfr, acc[k]);
                }
            }
        }
    }
    float *dst = partial + (int64_t)blockIdx.x * kNParams + kOffW2;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const int tap = wave + 4 * k;
        if (tap < 27) {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int ci = 8 * (v >> 2) + 4 * hi + (v & 3);
                dst[(tap * kCh + ci) * kCh + i] = acc[k][v];
            }
        }
    }
}

// dW1[t][ci][co] = sum_p feat[nbr(p, t)][ci] dpre1[p][co]; C rows m = t*4+ci
// (108, four blocks of 32, one per wave), columns co, K = p.
__global__ void __launch_bounds__(256) k_wgrad1(const uint16_t *__restrict__ feat,
                                                const uint16_t *__restrict__ dpre1, int64_t n_tiles,
                                                float *__restrict__ partial)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 31, hi = lane >> 5;
    const int m = wave * 32 + i, tap = m >> 2, ci = m & 3;
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint16_t *F = feat + t * kP * 4;
        const uint16_t *D = dpre1 + t * kP * kCh;
        for (int kb = 0; kb < kP / 16; ++kb) {
            const int p0 = kb * 16 + 8 * hi;
            uint32_t av[8], bv[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int q = m < 108 ? nbr(p0 + e, tap) : -1;
                av[e] = q >= 0 ? (uint32_t)F[q * 4 + ci] : 0u;
                bv[e] = (uint32_t)D[(p0 + e) * kCh + i];
            }
            const uint4 af = make_uint4(av[0] | (av[1] << 16), av[2] | (av[3] << 16), av[4] | (av[5] << 16),
                                        av[6] | (av[7] << 16));
            const uint4 bfr = make_uint4(bv[0] | (bv[1] << 16), bv[2] | (bv[3] << 16), bv[4] | (bv[5] << 16),
                                         bv[6] | (bv[7] << 16));
            acc = mfma(af, bfr, acc);
        }
    }
    float *dst = partial + (int64_t)blockIdx.x * kNParams + kOffW1;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = wave * 32 + 8 * (v >> 2) + 4 * hi + (v & 3);
        if (row < 108) dst[row * kCh + i] = acc[v];
    }
}

// dW3[t][ci] = sum_q act2[q][ci] dlogit[nbr(q, -t)] (C rows ci, columns t,
// K = q split over the 4 waves), and the bias gradients
// db1 = sum dpre1, db2 = sum dpre2, db3 = sum dlogit.
__global__ void __launch_bounds__(256) k_wgrad3_bias(const uint16_t *__restrict__ act2,
                                                     const float *__restrict__ dlogit,
                                                     const uint16_t *__restrict__ dpre1,
                                                     const uint16_t *__restrict__ dpre2, int64_t n_tiles,
                                                     float *__restrict__ partial)
{
    __shared__ float red[4][32 * 32 + 65];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 31, hi = lane >> 5;
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    float bsum = 0.0f;  // thread c < 32: db1[c]; 32..63: db2[c-32]; 64: db3 (per wave share)
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint16_t *A = act2 + t * kP * kCh;
        const float *D = dlogit + t * kP;
        for (int kb = wave; kb < kP / 16; kb += 4) {
            const int q0 = kb * 16 + 8 * hi;
            uint32_t av[8], bv[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                av[e] = (uint32_t)A[(q0 + e) * kCh + i];
                const int q = i < 27 ? nbr(q0 + e, 26 - i) : -1;
                bv[e] = q >= 0 ? f2bf(D[q]) : 0u;
            }
            const uint4 af = make_uint4(av[0] | (av[1] << 16), av[2] | (av[3] << 16), av[4] | (av[5] << 16),
                                        av[6] | (av[7] << 16));
            const uint4 bfr = make_uint4(bv[0] | (bv[1] << 16), bv[2] | (bv[3] << 16), bv[4] | (bv[5] << 16),
                                         bv[6] | (bv[7] << 16));
            acc = mfma(af, bfr, acc);
        }
        // bias sums: lane-owned channel over this wave's quarter of the visits
        const int c = lane;
        const uint16_t *src = c < 32 ? dpre1 + t * kP * kCh : dpre2 + t * kP * kCh;
        for (int p = wave; p < kP; p += 4) bsum += bf2f(src[p * kCh + (c & 31)]);
    }
    float db3 = 0.0f;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x)
        for (int p = threadIdx.x; p < kP; p += 256) db3 += dlogit[t * kP + p];
    db3 = wave_sum(db3);
#pragma unroll
    for (int v = 0; v < 16; ++v) red[wave][(8 * (v >> 2) + 4 * hi + (v & 3)) * 32 + i] = acc[v];
    red[wave][1024 + lane] = bsum;
    if (lane == 0) red[wave][1024 + 64] = db3;
    __syncthreads();
    float *dst = partial + (int64_t)blockIdx.x * kNParams;
    for (int e = threadIdx.x; e < 32 * 32; e += 256) {
        const int ci = e >> 5, tp = e & 31;
        const float s = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
        if (tp < 27) dst[kOffW3 + tp * kCh + ci] = s;
    }
    if (threadIdx.x < 64) {
        const int c = threadIdx.x;
        const float s = (red[0][1024 + c] + red[1][1024 + c]) + (red[2][1024 + c] + red[3][1024 + c]);
        dst[(c < 32 ? kOffB1 : kOffB2) + (c & 31)] = s;
    }
    if (threadIdx.x == 0)
        dst[kOffB3] = (red[0][1088] + red[1][1088]) + (red[2][1088] + red[3][1088]);
}

__global__ void k_reduce(const float *__restrict__ partial, int slots, float *__restrict__ grad)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kNParams) return;
    float s = 0.0f;
    for (int j = 0; j < slots; ++j) s += partial[(int64_t)j * kNParams + k];
    grad[k] = s;
}

__global__ void k_adam(float *__restrict__ th, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kNParams) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    th[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + eps);
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------

void launch_arrange(const float *theta, const Frags &f, hipStream_t s)
{
    hipLaunchKernelGGL(k_arrange, dim3(kK1 + 2 * kK2 + 2 * kK3), dim3(64), 0, s, theta, f);
}

void launch_features(const float2 *spec, int64_t n_tiles, uint16_t *feat, hipStream_t s)
{
    if (n_tiles > 0) hipLaunchKernelGGL(k_features, dim3((uint32_t)n_tiles), dim3(256), 0, s, spec, feat);
}

void launch_targets(const float2 *spec, const float2 *truth, int64_t n_tiles, float *w, float *num, hipStream_t s)
{
    if (n_tiles > 0) hipLaunchKernelGGL(k_targets, dim3((uint32_t)n_tiles), dim3(256), 0, s, spec, truth, w, num);
}

void launch_forward(const float *theta, const Frags &f, const Acts &a, int64_t n_tiles, float *mask, hipStream_t s)
{
    if (n_tiles <= 0) return;
    const dim3 grid((uint32_t)n_tiles), blk(256);
    hipLaunchKernelGGL(k_conv<kFwd1>, grid, blk, 0, s, (const void *)a.feat, f.w1, theta + kOffB1,
                       (const uint16_t *)nullptr, a.act1);
    hipLaunchKernelGGL(k_conv<kFwd2>, grid, blk, 0, s, (const void *)a.act1, f.w2, theta + kOffB2,
                       (const uint16_t *)nullptr, a.act2);
    hipLaunchKernelGGL(k_head, grid, blk, 0, s, a.act2, f.w3t, theta + kOffB3, a.logit, mask);
}

void launch_infer(const void *in, int64_t n_tiles, const float *theta, const Frags &f, float *mask, int cus,
                  hipStream_t s, bool paired)
{
    if (n_tiles <= 0) return;
    const int64_t units = paired ? (n_tiles + 1) / 2 : n_tiles;
    const int64_t g = std::min<int64_t>(units, 2 * std::max(1, cus));
    if (paired)
        hipLaunchKernelGGL(k_infer<true>, dim3((uint32_t)g), dim3(64 * kInferWaves), 0, s, in, n_tiles, theta, f,
                           mask);
    else
        hipLaunchKernelGGL(k_infer<false>, dim3((uint32_t)g), dim3(64 * kInferWaves), 0, s, in, n_tiles, theta, f,
                           mask);
}

void launch_backward(const float *theta, const Frags &f, const Acts &a, const float *mask, int64_t n_tiles,
                     float scale, float *partial, float *grad, hipStream_t s)
{
    (void)theta;
    if (n_tiles <= 0) return;
    const dim3 grid((uint32_t)n_tiles), blk(256);
    hipLaunchKernelGGL(k_loss, grid, blk, 0, s, mask, a.w, a.num, scale, a.dlogit, a.loss_tile);
    hipLaunchKernelGGL(k_conv<kBwd3>, grid, blk, 0, s, (const void *)a.dlogit, f.w3, (const float *)nullptr,
                       (const uint16_t *)a.act2, a.dpre2);
    hipLaunchKernelGGL(k_conv<kBwd2>, grid, blk, 0, s, (const void *)a.dpre2, f.w2t, (const float *)nullptr,
                       (const uint16_t *)a.act1, a.dpre1);
    const int slots = (int)(n_tiles < kWgradSlots ? n_tiles : kWgradSlots);
    hipLaunchKernelGGL(k_wgrad1, dim3(slots), blk, 0, s, a.feat, a.dpre1, n_tiles, partial);
    hipLaunchKernelGGL(k_wgrad2, dim3(slots), blk, 0, s, a.act1, a.dpre2, n_tiles, partial);
    hipLaunchKernelGGL(k_wgrad3_bias, dim3(slots), blk, 0, s, a.act2, a.dlogit, a.dpre1, a.dpre2, n_tiles,
                       partial);
    hipLaunchKernelGGL(k_reduce, dim3((kNParams + 255) / 256), dim3(256), 0, s, partial, slots, grad);
}

void launch_adam(float *theta, const float *grad, float *m, float *v, float lr, float beta1, float beta2, float eps,
                 int step, hipStream_t s)
{
    const float c1 = 1.0f / (1.0f - powf(beta1, (float)step));
    const float c2 = 1.0f / (1.0f - powf(beta2, (float)step));
    hipLaunchKernelGGL(k_adam, dim3((kNParams + 255) / 256), dim3(256), 0, s, theta, grad, m, v, lr, beta1, beta2,
                       eps, c1, c2);
}

}  // namespace mnet
