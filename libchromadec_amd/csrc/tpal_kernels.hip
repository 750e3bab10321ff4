// Transform-PAL on gfx950: tile transforms, carrier-mirror mask, overlap-add,
// PAL demodulation and output conversion.
//
// The 3D tile transform follows TransformPal3D::filterFields
// (src/decoders/palcolour/transform_pal_3d.cpp:123-333): a raised-cosine
// windowed 8x32x16 tile, forward real FFT, keep a bin and its mirror about
// the U/V carrier when their magnitudes agree within the threshold, inverse
// FFT, overlap-add at half-tile hops.  It is computed here in FP32 with an
// exact reformulation that skips all work the output cannot see:
//   * black16bIre is subtracted from every sample.  The window's spectrum
//     occupies bins {0, +-1} per axis and the mask passes only x bins 2..6,
//     so this changes nothing the mask lets through, while the black-filled
//     lines of the other field become exact zeros.
//   * Each z plane holds data on one line parity p(z) only, so its 32-point
//     vertical transform is a 16-point transform of the data lines times a
//     parity twiddle.  The inverse likewise folds 32 -> 16 bins per plane
//     and yields only the lines of that plane's field.
//   * Only x bins 2..6 are formed (direct windowed DFT of 16 samples) and
//     only the output planes the frame lattice uses are inverted.
// The magnitude comparison runs in FP32; a comparison whose outcome FP32
// rounding could change (margin below a bound on the FP32 error of the two
// bins) is decided again from FP64 bins summed directly from the samples,
// so the keep/discard decisions match an FP64 evaluation.
//
// The demodulator follows PalColour::detectBurst and
// PalColour::decodeLine<double, true> (src/decoders/palcolour/palcolour.cpp:
// 344-409, 456-646) and the output stage OutputWriter::convertToFloat /
// convertLine (src/output/output_writer.cpp:402-486, 670-727).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "mask_feature.h"
#include "tpal_internal.h"

namespace tpal {

// ---------------------------------------------------------------------------
// Small complex helper and compile-time twiddles (multiples of 2*pi/32).
// ---------------------------------------------------------------------------

struct c32 {
    float x, y;
};

// Global-memory accesses through vector-memory instructions.  A generic
// pointer gives flat instructions, which also count as LDS traffic, so every
// wait for LDS (before each barrier) would wait for them too.  (Native vector
// types: a HIP vector struct is copied as bytes and loses the address space.)
typedef float gf32x4 __attribute__((ext_vector_type(4)));
typedef unsigned gu32x4 __attribute__((ext_vector_type(4)));
typedef unsigned gu32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 gload_f4(const float *p)
{
    const gf32x4 v = *(const __attribute__((address_space(1))) gf32x4 *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 gload_u4(const void *p)
{
    const gu32x4 v = *(const __attribute__((address_space(1))) gu32x4 *)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 gload_u2(const uint16_t *p)
{
    const gu32x2 v = *(const __attribute__((address_space(1))) gu32x2 *)p;
    return make_uint2(v.x, v.y);
}
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) uint16_t gu16;
__device__ __forceinline__ void gstore_f4(float *p, float a, float b, float c, float d)
{
    gf32x4 v = {a, b, c, d};
    *(__attribute__((address_space(1))) gf32x4 *)p = v;
}
typedef float gf32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void gstore_f2(float *p, float a, float b)
{
    gf32x2 v = {a, b};
    *(__attribute__((address_space(1))) gf32x2 *)p = v;
}
__device__ __forceinline__ c32 cadd(c32 a, c32 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ c32 csub(c32 a, c32 b) { return {a.x - b.x, a.y - b.y}; }

__device__ constexpr float kCosOct[9] = {
    1.0f, 0.98078528040323043f, 0.92387953251128674f, 0.83146961230254524f,
    0.70710678118654757f, 0.55557023301960229f, 0.38268343236508984f,
    0.19509032201612833f, 0.0f};

__device__ __forceinline__ float cos32(int k)
{
    k &= 31;
    return k <= 8 ? kCosOct[k] : (k <= 16 ? -kCosOct[16 - k] : (k <= 24 ? -kCosOct[k - 16] : kCosOct[32 - k]));
}
__device__ __forceinline__ float sin32(int k) { return cos32(k + 24); }

// v * exp(SIGN * 2*pi*i * t / 32), with the quarter turns done exactly.
template <int SIGN>
__device__ __forceinline__ c32 twiddle(c32 b, int t)
{
    t &= 31;
    if (t == 0) return b;
    if (t == 8) return SIGN < 0 ? c32{b.y, -b.x} : c32{-b.y, b.x};
    if (t == 16) return c32{-b.x, -b.y};
    if (t == 24) return SIGN < 0 ? c32{-b.y, b.x} : c32{b.y, -b.x};
    const float c = cos32(t), s = SIGN * sin32(t);
    return c32{b.x * c - b.y * s, b.x * s + b.y * c};
}

template <int N>
__device__ __forceinline__ void bitrev_perm(c32 (&v)[N])
{
#pragma unroll
    for (int i = 0; i < N; ++i) {
        int j = 0;
#pragma unroll
        for (int b = 0; (1 << b) < N; ++b)
            if (i & (1 << b)) j |= (N >> 1) >> b;
        if (i < j) {
            c32 t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
}

// In-register radix-2 FFT, unnormalised.  SIGN = -1 forward, +1 inverse.
template <int N, int SIGN>
__device__ __forceinline__ void fft(c32 (&v)[N])
{
    bitrev_perm(v);
#pragma unroll
    for (int len = 2; len <= N; len <<= 1) {
#pragma unroll
        for (int i = 0; i < N; i += len) {
#pragma unroll
            for (int k = 0; k < len / 2; ++k) {
                const c32 a = v[i + k];
                const c32 w = twiddle<SIGN>(v[i + k + len / 2], k * (32 / len));
                v[i + k] = cadd(a, w);
                v[i + k + len / 2] = csub(a, w);
            }
        }
    }
}

// Constant tables, filled once per device by the host (tpal_device.cpp).
__constant__ Tables c_tab;

// What a tile-row load reads when its row is black (outside the tile's lines
// or an absent field): every lane loads, so that the loads sit on every path
// and the compiler's wait counts stay exact; the row's weight is 0.
__device__ __align__(16) uint16_t g_no_row[16];

// Error bound for FP32 bins: |bin error| <= kErrK * 2^-24 * L1, with L1 the
// sum of |windowed, black-subtracted samples| of the tile.  A worst-case
// rounding analysis of the folded x DFT, the radix-2 stages and the twiddles
// gives about 34 u L1 (3D) and 45 u L1 (2D, for 16-term sums per x bin; the
// folded x DFT now used rounds fewer times on every path); the largest error measured over
// 384 000 bins of synthetic PAL tiles (noise, zone plates, bars) was
// 0.69 u L1.  The bound used keeps a factor of >10 over every measured error.
constexpr float kErrK = 8.0f;    // 3D tile
constexpr float kErrK2 = 12.0f;  // 2D tile (32-sample x DFT)
constexpr float kU = 5.9604645e-08f;  // 2^-24

// max / min of |x| and |y| in one instruction each.  The compiler's fmaxf
// first canonicalises each input (IEEE mode), which costs two more; the bins
// here are never NaN, so the canonicalisation changes nothing.
__device__ __forceinline__ float absmax(float x, float y)
{
    float r;
    asm("v_max_f32 %0, |%1|, |%2|" : "=v"(r) : "v"(x), "v"(y));
    return r;
}
__device__ __forceinline__ float absmin(float x, float y)
{
    float r;
    asm("v_min_f32 %0, |%1|, |%2|" : "=v"(r) : "v"(x), "v"(y));
    return r;
}

// Keep/discard decision with an FP32 uncertainty test
// (transform_pal_3d.cpp:313-327: discard when |a|^2 < |b|^2 t^2 or
// |b|^2 < |a|^2 t^2): keep, discard, or unsure when FP32 error `e` on the
// bins could change the outcome.  |m_hat - m| <= 2|a|e + e^2, plus the
// rounding of the products themselves.  The callers' bound e is zero only
// for a tile whose windowed samples are all zero (a 2D padding step past its
// strip, or a tile of black): every bin is then exactly zero and the pair is
// kept (0 < 0 is false), as an FP64 evaluation decides; the callers keep such
// tiles' pairs without the test.
// The terms of the bounds that depend only on the tile's e and the
// threshold th: d1 = 2e na + 2e th nb + e^2 (1 + th) + 4u ma + 4u th mb, and
// d2 the same with a and b swapped.
struct DecideConst {
    float th, c2e, c2eth, e2th1, u4th;
};
__device__ __forceinline__ DecideConst decide_const(float e, float th)
{
    return DecideConst{th, 2.0f * e, 2.0f * e * th, e * e * (1.0f + th), 4.0f * kU * th};
}

// The decision as wave ballots (taken on the comparisons themselves, so that
// no per-lane bool is materialised): keep and unsure lane masks.
__device__ __forceinline__ void decide32_ballots(c32 a, c32 b, const DecideConst &q, uint64_t &keep,
                                                 uint64_t &unsure)
{
    const float ma = fmaf(a.x, a.x, a.y * a.y);
    const float mb = fmaf(b.x, b.x, b.y * b.y);
    // |a| <= max(|a.x|,|a.y|) + (sqrt 2 - 1) min(|a.x|,|a.y|)
    const float na = fmaf(0.41421357f, absmin(a.x, a.y), absmax(a.x, a.y));
    const float nb = fmaf(0.41421357f, absmin(b.x, b.y), absmax(b.x, b.y));
    const float g1 = fmaf(-q.th, mb, ma), g2 = fmaf(-q.th, ma, mb);
    const float d1 = fmaf(q.c2e, na, fmaf(q.c2eth, nb, fmaf(4.0f * kU, ma, fmaf(q.u4th, mb, q.e2th1))));
    const float d2 = fmaf(q.c2e, nb, fmaf(q.c2eth, na, fmaf(4.0f * kU, mb, fmaf(q.u4th, ma, q.e2th1))));
    // one ballot per comparison, combined as scalar masks
    unsure = __builtin_amdgcn_ballot_w64(fabsf(g1) <= d1) | __builtin_amdgcn_ballot_w64(fabsf(g2) <= d2);
    const uint64_t discard = __builtin_amdgcn_ballot_w64(fminf(g1, g2) < 0.0f);
    keep = __builtin_amdgcn_read_exec() & ~(unsure | discard);
}

// cos / sin (2 pi j / 16) with the exact values at multiples of pi/2.
constexpr double kC1 = 0.9238795325112867, kS1 = 0.3826834323650898, kR2 = 0.7071067811865476;
constexpr double kCos16[16] = {1.0, kC1, kR2, kS1, 0.0, -kS1, -kR2, -kC1, -1.0, -kC1, -kR2, -kS1, 0.0, kS1, kR2, kC1};
constexpr double kSin16[16] = {0.0, kS1, kR2, kC1, 1.0, kC1, kR2, kS1, 0.0, -kS1, -kR2, -kC1, -1.0, -kC1, -kR2, -kS1};

// The value of the neighbouring lane of a pair (lane ^ 1): one DPP move,
// quad_perm 1,0,3,2.
__device__ __forceinline__ float swap_pair(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
}

// One DPP move of a double (its two halves); rows or banks outside the masks
// read 0.  With every row and bank enabled (the permutations used here read
// a lane of every row) no lane keeps its old value, so none is zeroed first.
template <int CTRL, int ROWMASK = 0xF, int BANKMASK = 0xF>
__device__ __forceinline__ double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    int lo, hi;
    if constexpr (ROWMASK == 0xF && BANKMASK == 0xF) {
        lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
        hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    } else {
        lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWMASK, BANKMASK, false);
        hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, BANKMASK, false);
    }
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// One DPP move of a float, as dpp_d.
template <int CTRL, int ROWMASK = 0xF, int BANKMASK = 0xF>
__device__ __forceinline__ float dpp_f(float v)
{
    if constexpr (ROWMASK == 0xF && BANKMASK == 0xF)
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
    else
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, BANKMASK, false));
}

// The wave's total of v, the same in every lane (a scalar): DPP sums into
// lane 63, as wave_sum4_to63 below, then that lane read out.
__device__ __forceinline__ float wave_total(float v)
{
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    v += dpp_f<0x141>(v);
    v += dpp_f<0x140>(v);
    v += dpp_f<0x142, 0xA>(v);
    v += dpp_f<0x143, 0xC>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Four wave sums into lane 63 with DPP moves, no LDS round trips: pairs
// (quad_perm 1,0,3,2), quads (2,3,0,1), half rows (row_half_mirror), rows
// (row_mirror), then row 1 += row 0 and row 3 += row 2 (row_bcast:15), and
// row 3 += rows 0-1 (row_bcast:31).  Other lanes hold partial sums.
__device__ __forceinline__ void wave_sum4_to63(double &a, double &b, double &c, double &d)
{
    a += dpp_d<0xB1>(a); b += dpp_d<0xB1>(b); c += dpp_d<0xB1>(c); d += dpp_d<0xB1>(d);
    a += dpp_d<0x4E>(a); b += dpp_d<0x4E>(b); c += dpp_d<0x4E>(c); d += dpp_d<0x4E>(d);
    a += dpp_d<0x141>(a); b += dpp_d<0x141>(b); c += dpp_d<0x141>(c); d += dpp_d<0x141>(d);
    a += dpp_d<0x140>(a); b += dpp_d<0x140>(b); c += dpp_d<0x140>(c); d += dpp_d<0x140>(d);
    a += dpp_d<0x142, 0xA>(a); b += dpp_d<0x142, 0xA>(b); c += dpp_d<0x142, 0xA>(c); d += dpp_d<0x142, 0xA>(d);
    a += dpp_d<0x143, 0xC>(a); b += dpp_d<0x143, 0xC>(b); c += dpp_d<0x143, 0xC>(c); d += dpp_d<0x143, 0xC>(d);
}

__device__ __forceinline__ int decide64(double ar, double ai, double br, double bi, double th)
{
    const double ma = ar * ar + ai * ai, mb = br * br + bi * bi;
    return (ma < mb * th || mb < ma * th) ? 0 : 1;
}

// ---------------------------------------------------------------------------
// 3D tile kernel.  One 64-lane workgroup = one strip of 4-24 (+1 leading)
// tiles of one tile row of one slab.
// ---------------------------------------------------------------------------

// LDS layout.  Rows are padded so that the lanes of a wave, which walk
// different rows in step, hit different banks: A rows are 17 complex
// (136 B) apart, S columns 9 complex (72 B) apart, and each k block of S is
// shifted by 2 complex.  A (x-DFT rows / inverse-y output) and S (spectrum)
// are never live at the same time, so they share storage.
constexpr int kApad = 17;
constexpr int kSpad = 9;

// HALF: the half form stores only the spectrum rows ky < 16, so its k blocks
// are 16 rows (and the workgroup needs 7 KB instead of 12.5).
template <bool HALF>
struct Lds3 {
    static constexpr int kSk = (HALF ? 16 : 32) * kSpad + 2;   // k block stride of S
    union {
        c32 A[40][kApad];     // forward: row z*5+k holds the 16 data lines m
                              // inverse: row zs*5+k holds the 16 output lines
        c32 S[5 * kSk];       // spectrum S[k*kSk + ky*kSpad + kz]
    } u;
    uint32_t keep[40];        // word k*8+kz, bit ky: bin (x = k+2, ky, kz) kept
    // mask step words, one 64-bit ballot per step (x, z < 4): word pw(x-2, z)
    // (below), bit y: visit (z, y, x) keeps its pair / is undecidable in FP32
    alignas(8) uint32_t bal[24];
    alignas(8) uint32_t flg[24];
    double cs[32], sn[32];    // cos/sin(2 pi j / 32) for the FP64 re-check
    double wz[8], wy[32];     // its z and y window factors
};

// The soft-mask modes hold the tile's mask as well: one value per visit, or
// per visit of planes z < 4 in the paired form.
template <bool HALF, int NM>
struct Lds3M : Lds3<HALF> {
    float m[NM];
};
// The masked inverse of the paired learned path stages each lane's columns
// of three tiles and stores whole 128-byte lines with the fourth:
// [tile % 4][half][lane].
template <class B>
struct LdsStaged : B {
    float4 stage[3][2][64];
};
template <int MODE, bool HALF> struct LdsFor { using T = Lds3<HALF>; };
template <bool HALF> struct LdsFor<kTileApply, HALF> { using T = Lds3M<HALF, 768>; };
template <bool HALF> struct LdsFor<kTileApplyFwd, HALF> { using T = LdsStaged<Lds3M<HALF, 384>>; };

template <class LT>
__device__ __forceinline__ c32 &Sat(LT &L, int k, int ky, int kz) { return L.u.S[k * LT::kSk + ky * kSpad + kz]; }

// Visit whose mask value applies to bin (x = k+2, ky, kz): its own visit for
// x = 2..4, the visit it is the carrier-mirror partner of for x = 5, 6.
__device__ __forceinline__ int mask_visit(int k, int ky, int kz)
{
    if (k <= 2) return (kz * 32 + ky) * 3 + k;
    return (((10 - kz) & 7) * 32 + ((40 - ky) & 31)) * 3 + (4 - k);
}

// Word of plane (k = x - 2, z) in the mask step arrays: the two planes a step
// decides, (k, z) and (k, z + 4), are one 64-bit word.
__device__ __forceinline__ constexpr int pw(int k, int z) { return (k * 4 + (z & 3)) * 2 + (z >> 2); }

// Bit y of w moved to bit (8 - y) mod 32: the mirror row yr = (40 - y) % 32
// of transform_pal_3d.cpp:291, applied to a word of 32 bins along ky.
__device__ __forceinline__ uint32_t mirror_rows(uint32_t w)
{
    const uint32_t r = __builtin_bitreverse32(w);  // bit y -> 31 - y
    return (r << 9) | (r >> 23);                   // + 9 (mod 32)
}

// acc + x * c for a coefficient known at compile time after unrolling; the
// exact 0 / +-1 cases cost nothing.
__device__ __forceinline__ float cmac(float acc, float x, float c)
{
    if (c == 0.0f) return acc;
    if (c == 1.0f) return acc + x;
    if (c == -1.0f) return acc - x;
    return fmaf(x, c, acc);
}

// 16 samples (two uint4 of packed u16) minus black, as floats.
__device__ __forceinline__ void unpack16(const uint4 u0, const uint4 u1, float black, float (&s)[16])
{
    const uint32_t w[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        s[2 * i] = (float)(w[i] & 0xFFFFu) - black;
        s[2 * i + 1] = (float)(w[i] >> 16) - black;
    }
}

// Windowed 16-sample real DFT to bins 2..6 (X[j] = bin j+2), folded by the
// half-period symmetries of the bins: even bins see u[n] + u[n+8], odd bins
// u[n] - u[n+8]; bins 2 and 6 fold again over n+4.  Returns sum |u[n]|.
__device__ __forceinline__ float xdft16(const float (&s)[16], c32 (&X)[5])
{
    float u[16], l1 = 0.0f;
#pragma unroll
    for (int n = 0; n < 16; ++n) {
        u[n] = s[n] * c_tab.wx[n];
        l1 += fabsf(u[n]);
    }
    float e[8], o[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        e[n] = u[n] + u[n + 8];
        o[n] = u[n] - u[n + 8];
    }
    // bin 4: sum e[n] (-i)^n
    X[2] = c32{(e[0] - e[2]) + (e[4] - e[6]), (e[3] - e[1]) + (e[7] - e[5])};
    // bins 2 and 6: g[n] = e[n] - e[n+4]
    const float g0 = e[0] - e[4], g1 = e[1] - e[5], g2 = e[2] - e[6], g3 = e[3] - e[7];
    const float r = 0.70710678118654752f;
    const float a = (g1 - g3) * r, b = (g1 + g3) * r;
    X[0] = c32{g0 + a, -(g2 + b)};
    X[4] = c32{g0 - a, g2 - b};
    // bins 3 and 5: with t = 2 pi 3 n / 16, bin 5's twiddle is (-1)^n times
    // the conjugate of bin 3's, so both come from the even-n and odd-n sums
    // of o[n] cos t and -o[n] sin t
    float ce = 0.0f, co = 0.0f, se = 0.0f, so = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; n += 2) {
        ce = cmac(ce, o[n], cos32(6 * n));
        se = cmac(se, o[n], -sin32(6 * n));
        co = cmac(co, o[n + 1], cos32(6 * (n + 1)));
        so = cmac(so, o[n + 1], -sin32(6 * (n + 1)));
    }
    X[1] = c32{ce + co, se + so};
    X[3] = c32{ce - co, so - se};
    return l1;
}

// c2r over bins 2..6 (C[j] = bin j+2, pre-scaled): out[n] = 2 Re sum C e^{+i..}.
// Even bins repeat every 8 samples, odd bins change sign.
__device__ __forceinline__ void xidft16(const c32 (&C)[5], float (&out)[16])
{
    // Sample n = 0 and 4 (quarter-turn twiddles), then n = 1..3 with 8 - n:
    // at t = 2 pi k n / 16, sample 8 - n has cos (-1)^k cos t and sin
    // -(-1)^k sin t, so both come from the even-k and odd-k sums of
    // C.x cos t and C.y sin t.  Sample n + 8 is E - O.
#pragma unroll
    for (int n = 0; n <= 4; n += 4) {
        float E = 0.0f, O = 0.0f;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const int k = j + 2;
            const float c = cos32(2 * k * n), s = sin32(2 * k * n);
            if (k & 1) {
                O = cmac(O, C[j].x, c);
                O = cmac(O, C[j].y, -s);
            } else {
                E = cmac(E, C[j].x, c);
                E = cmac(E, C[j].y, -s);
            }
        }
        out[n] = E + O;
        out[n + 8] = E - O;
    }
#pragma unroll
    for (int n = 1; n < 4; ++n) {
        float pc = 0.0f, ps = 0.0f, qc = 0.0f, qs = 0.0f;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const int k = j + 2;
            const float c = cos32(2 * k * n), s = sin32(2 * k * n);
            if (k & 1) {
                qc = cmac(qc, C[j].x, c);
                qs = cmac(qs, C[j].y, s);
            } else {
                pc = cmac(pc, C[j].x, c);
                ps = cmac(ps, C[j].y, s);
            }
        }
        const float En = pc - ps, Em = pc + ps;       // E at n and 8 - n
        const float On = qc - qs, Om = -(qc + qs);    // O at n and 8 - n
        out[n] = En + On;
        out[n + 8] = En - On;
        out[8 - n] = Em + Om;
        out[16 - n] = Em - Om;
    }
}

// Per-phase clock stamps of the plain-mode 3D tile kernel, compiled in only
// with -DCHD_TILE_PHASES (make -C libchromadec_amd phases; bench.py
// --tile-phases reads them): lane i of a workgroup accumulates the shader
// clocks its wave spends in phase i (0 forward x, 1 vertical, 2 temporal,
// 3 mask steps, 4 keep words, 5 FP64 re-decision, 6 inverse z, 7 inverse y,
// 8 inverse x and overlap-add, 9 tile setup) and the counts 16 (tiles with a
// re-decision), 17 (flagged words) and 18 (tiles), and adds them to
// g_tile_phases at exit.  The stamps wait for nothing but the clock read.
#ifdef CHD_TILE_PHASES
__device__ unsigned long long g_tile_phases[32];
#define CHD_PHASE_INIT() uint64_t ph_acc = 0, ph_prev = __builtin_amdgcn_s_memtime()
#define CHD_PHASE(i)                                                                                          \
    do {                                                                                                      \
        if constexpr (MODE == kTilePlain) {                                                                   \
            const uint64_t t_ = __builtin_amdgcn_s_memtime();                                                 \
            ph_acc += (lane0 == (i)) ? (t_ - ph_prev) : 0ull;                                                 \
            ph_prev = t_;                                                                                     \
        }                                                                                                     \
    } while (0)
#define CHD_PHASE_COUNT(i, n)                                                                                 \
    do {                                                                                                      \
        if constexpr (MODE == kTilePlain) ph_acc += (lane0 == (i)) ? (unsigned long long)(n) : 0ull;          \
    } while (0)
#define CHD_PHASE_FLUSH()                                                                                     \
    do {                                                                                                      \
        if constexpr (MODE == kTilePlain)                                                                     \
            if (lane0 < 32) atomicAdd(&g_tile_phases[lane0], ph_acc);                                         \
    } while (0)
#else
#define CHD_PHASE_INIT() \
    do {                 \
    } while (0)
#define CHD_PHASE(i) \
    do {             \
    } while (0)
#define CHD_PHASE_COUNT(i, n) \
    do {                      \
    } while (0)
#define CHD_PHASE_FLUSH() \
    do {                  \
    } while (0)
#endif

// MODE kTilePlain: the reference's hard mask.  kTileSpectra: stop after the
// forward transform and write each tile's bins x = 2..6 to `spec`
// ([tile][k][ky][kz], tile = (job * n_ty3 + ty) * n_tx3 + tx).
// kTileApply: read those bins back, weight each by the soft mask of its
// visit (`mask`, [tile][768]) and run the inverse and overlap-add.
// The paired learned path (HALF only): kTileFeatures stops after the forward
// transform and writes the magnitude feature of each bin of rows ky < 16
// (`spec` as floats, [tile][k][ky < 16][kz], 640 per tile), all the paired
// inference reads; kTileApplyFwd computes the forward transform again from
// the composite (bit for bit the same bins) and applies the soft mask.
// Work item = one strip of `strip` tiles of one tile row of one slab.  The
// grid is either one workgroup per item (PERSIST false) or persistent (a
// multiple of 8 workgroups, each looping over items).  HALF (plain mode):
// every slab alternates in line parity and one threshold applies to every
// visit.
template <int MODE, bool PERSIST, bool HALF>
__global__ void __launch_bounds__(64, 3)   // 3 waves per SIMD: at most 168 VGPRs
tpal3d_tiles(const Geom g, const SlabJob *__restrict__ jobs, int n_jobs, int strip_len, const float *__restrict__ thr32,
             const double *__restrict__ thr64, float2 *__restrict__ spec, const float *__restrict__ mask,
             unsigned *__restrict__ wq)
{
    __shared__ typename LdsFor<MODE, HALF>::T L;
    constexpr int kSk = LdsFor<MODE, HALF>::T::kSk;
    const int lane0 = threadIdx.x;
    // the c2r scale of the inverse (2 / 4096; the half form's fold doubles it)
    constexpr float kInvScale = (HALF ? 4.0f : 2.0f) / 4096.0f;
    static_assert(MODE == kTilePlain || HALF == (MODE == kTileFeatures || MODE == kTileApplyFwd),
                  "the paired modes are half-form only, the stored-spectra modes full-form only");
    // forward transform from the composite / soft mask in the inverse /
    // inverse scale folded into the forward row weights
    constexpr bool kForward = MODE != kTileApply;
    constexpr bool kSoft = MODE == kTileApply || MODE == kTileApplyFwd;
    constexpr bool kFoldScale = MODE == kTilePlain || MODE == kTileApplyFwd;
    // overlap-add stores of whole lines through an LDS stage
    constexpr bool kStaged = MODE == kTileApplyFwd;
    CHD_PHASE_INIT();
    if (lane0 < 32) {
        L.cs[lane0] = c_tab.cos32_d[lane0];
        L.sn[lane0] = c_tab.sin32_d[lane0];
        L.wy[lane0] = c_tab.wy_d[lane0];
        if (lane0 < 8) L.wz[lane0] = c_tab.wz_d[lane0];
    }
    // Workgroups are dispatched round-robin over the 8 XCDs, each with its
    // own L2.  Give each XCD a contiguous share of the items, ordered slab
    // fastest, then tile row, then strip: the workgroups resident on one XCD
    // then cover the same few field columns of neighbouring slabs (which
    // share 6 of their 8 fields) and neighbouring tile rows (which share 8
    // of their 16 lines), so each field line is fetched into that L2 about
    // once instead of once per tile.
    const int n_strips = (g.n_tx3 + strip_len - 1) / strip_len;
    const int total = n_jobs * g.n_ty3 * n_strips;
    const int G = gridDim.x;
    const int xcd = blockIdx.x & 7, q = total >> 3, r = total & 7;
    const int lo = xcd * q + min(xcd, r), hi = lo + q + (xcd < r ? 1 : 0);
    // One workgroup per item: item lo + (its rank on the XCD).  Persistent:
    // the workgroups of an XCD take that XCD's items in order from its own
    // counter wq[xcd], so uneven items balance as the hardware's dispatch of
    // one workgroup per item would; wq[8] counts workgroups that ran out, and
    // the last of them resets all nine counters for the next launch (every
    // other workgroup has made its last fetch by then).
    auto fetch = [&]() -> int {
        if constexpr (PERSIST) {
            unsigned k = 0;
            if (lane0 == 0) k = atomicAdd(&wq[xcd], 1u);
            return lo + (int)__builtin_amdgcn_readfirstlane(k);
        } else {
            return lo + (blockIdx.x >> 3);
        }
    };
    for (int w = fetch(); w < hi; w = PERSIST ? fetch() : hi) {
    // A persistent workgroup recomputes its per-lane setup for each item
    // from an opaque copy of the lane index: hoisted out of the item loop,
    // that setup would stay live through every tile and spill.
    int lane = lane0;
    if constexpr (PERSIST) asm volatile("" : "+v"(lane));
    const int job_id = w % n_jobs;
    const int rest = w / n_jobs;
    const int ty = rest % g.n_ty3;
    const int strip = rest / g.n_ty3;
    const SlabJob &job = jobs[job_id];

    const int tileY = g.first - kYT / 2 + (kYT / 2) * ty;
    const int startY = max(g.first - tileY, 0);
    const int endY = min(g.last - tileY + 1, kYT);
    int par[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) par[z] = (job.offset[z] - tileY) & 1;
    // the same as a scalar bit mask (bit z: plane z holds odd lines), for
    // uniform branches
    int pbits = 0;
#pragma unroll
    for (int z = 0; z < 8; ++z) pbits |= par[z] << z;
    pbits = __builtin_amdgcn_readfirstlane(pbits);

    const int tx_own0 = strip * strip_len;
    const int tx_end = min(tx_own0 + strip_len, g.n_tx3);
    // the leading tile completes the first owned columns; spectra need none
    const int tx_begin = (MODE == kTileSpectra || MODE == kTileFeatures) ? tx_own0 : max(tx_own0 - 1, 0);

    // forward-stage role: rows (z, m) = lane and lane + 64
    const uint16_t *rowp[2];
    float roww[2];
    int rz[2], ry[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int r = lane + 64 * h;
        rz[h] = r >> 4;
        ry[h] = 2 * (r & 15) + par[rz[h]];
        const uint16_t *fp = job.field[rz[h]];
        const bool ok = fp != nullptr && ry[h] >= startY && ry[h] < endY;
        rowp[h] = ok ? fp + (int64_t)((tileY + ry[h]) >> 1) * g.pitch + g.lmargin : nullptr;
        // the modes that run the inverse after their own forward transform
        // fold its scale (a power of two, so the products round the same)
        // into the forward row weight.  A black row (outside the tile's
        // lines, or an absent field) loads a stand-in and gets weight 0: its
        // bins and its share of the error bound are then exactly 0, as for
        // black samples.
        roww[h] = ok ? c_tab.wz[rz[h]] * c_tab.wy[ry[h]] * (kFoldScale ? kInvScale : 1.0f) : 0.0f;
    }
    // vertical-stage roles: forward lanes (z, k) = lane / 5, inverse lanes
    // (zs, k) = lane / 5 with z = 0,1,4,5
    const bool fy_odd = par[min(lane / 5, 7)] != 0;
    // Line parity alternating with z (consecutive fields of a normal
    // capture): row ky+16 of plane z is (-1)^par[z] = sgn (-1)^z times row ky,
    // so after the temporal DFT T[ky+16][kz] = sgn T[ky][(kz+4) & 7] and only
    // rows 0..15 need the vertical stores and the temporal DFT.
    // (the launcher checked the parity.  The features mode then writes the
    // magnitudes of rows ky < 16 for the paired inference, and the
    // apply-forward mode reads its masks, which repeat at (z + 4, y + 16).)
    constexpr bool kHalf = HALF;
    const bool alt = kHalf || pbits == 0xAA || pbits == 0x55;
    const float sgn = par[0] ? -1.0f : 1.0f;
    // With alternating parity and one threshold for every visit, the rows
    // ky >= 16 follow from the rows ky < 16 all the way through:
    //   * visit (z + 4, y + 16, x) compares the magnitudes of visit (z, y, x)
    //     (its bins and their mirrors are +-the same FP32 values), so its
    //     keep decision is the same, and the FP64 re-decision (twiddles
    //     antisymmetric over half a period) agrees too: keep word (x, z + 4)
    //     is keep word (x, z) rotated by 16 rows;
    //   * the masked inverse over z then gives T[ky + 16][z] =
    //     sgn (-1)^z T[ky][z], and for every output plane the vertical fold
    //     T[ky] + (-1)^p T[ky + 16] is 2 T[ky].
    // So only rows ky < 16 are stored, only planes z < 4 decided and only
    // rows ky < 16 inverted over z.  (A kernel of its own: the general form
    // beside it would raise the register count of both.)
    constexpr bool half = kHalf;
    // exp(-2 pi i ky / 32) for this lane's temporal-stage ky (lane & 15; exact
    // at the quarter turns)
    const c32 ptw = {cos32(lane & 15), -sin32(lane & 15)};
    // exp(+2 pi i ky / 32) for this lane's inverse-z ky (lane & 31)
    const c32 itw = {cos32(lane & 31), sin32(lane & 31)};
    // mask-stage role: bins y = lane & 31; lanes 0-31 visit x = 2 (all z) and
    // x = 3 (z < 4), lanes 32-63 visit x = 3 (z >= 4) and x = 4 (all z)
    const int vy = lane & 31, vhalf = lane >> 5;

    // inverse-stage role: output plane zsel (z = 0,1,4,5), line om
    const int zsel = lane >> 4, om = lane & 15;
    const int oz = (zsel < 2) ? zsel : zsel + 2;
    const int opar = (job.offset[oz] - tileY) & 1;
    const int oy = 2 * om + opar;
    const bool orow_ok = (oy >= startY) && (oy < endY);
    float *orow = (MODE == kTileSpectra || MODE == kTileFeatures) ? nullptr
                                       : job.partial + (int64_t)(zsel * 2 + (ty & 1)) * g.part_plane +
                                             (int64_t)((tileY + oy) >> 1) * g.cpitch;
    // where a lane with no output columns stores: columns 0..31 of its
    // plane's first row, the same for every lane of the plane, so that these
    // stores meet in one line of the cache (the demodulators read the active
    // columns, from kColOrigin on, and the tile stores start at
    // kColOrigin - 8)
    float *orow_none = (MODE == kTileSpectra || MODE == kTileFeatures)
                           ? nullptr
                           : job.partial + (int64_t)(zsel * 2 + (ty & 1)) * g.part_plane;
    float carry[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) carry[i] = 0.0f;

    // apply mode: the next tile's spectrum and mask are loaded into registers
    // while the current tile is inverted (its loads would otherwise wait on
    // HBM once per tile).  The half form reads the compact spectra of the
    // paired path, rows ky < 16 only ([tile][k][ky < 16][kz], 640 bins).
    // soft masks: one per visit ([tile][768]), or in the paired form one per
    // visit of planes z < 4 ([tile][384]; visit (z + 4, y + 16) repeats
    // visit (z, y))
    constexpr int kSpecBins = 1280;
    constexpr int kSpecPerLane = (MODE == kTileApply) ? kSpecBins / 64 : 1;
    constexpr int kMaskVisits = MODE == kTileApplyFwd ? 384 : 768;
    constexpr int kMaskPerLane = kSoft ? kMaskVisits / 64 : 1;
    float2 pre_s[kSpecPerLane];
    float pre_m[kMaskPerLane];
    auto prefetch = [&](int64_t t) {
        if constexpr (MODE == kTileApply) {
#pragma unroll
            for (int j = 0; j < kSpecPerLane; ++j) pre_s[j] = spec[t * kSpecBins + lane + 64 * j];
        }
        if constexpr (kSoft) {
#pragma unroll
            for (int j = 0; j < kMaskPerLane; ++j) pre_m[j] = mask[t * kMaskVisits + lane + 64 * j];
        }
    };

    // This lane's two field rows of a tile (16 samples each); a row outside
    // the tile's lines reads as black (zero after black subtraction).  They
    // are loaded during the previous tile's inverse, and serve the forward
    // x transform and the FP64 re-decisions.
    uint4 rows[2][2];
    auto load_rows = [&](int txl) {
        const int tX = g.avs - kXT / 2 + (kXT / 2) * txl;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            // (a black row reads a stand-in, weighted 0 in the transform)
            const uint16_t *p = rowp[h] != nullptr ? rowp[h] + tX : g_no_row;
            rows[h][0] = gload_u4(p);
            rows[h][1] = gload_u4(p + 8);
        }
    };
    if constexpr (kForward) {
        load_rows(tx_begin);
        // The first tile's rows are waited for here (s_waitcnt vmcnt(0)):
        // the loop entry then has no row load in flight, and the loop waits
        // for each next tile's rows by count, with the tile's stores (and the
        // next mask) still in flight.
        __builtin_amdgcn_s_waitcnt(0x0F70);
    }
    __builtin_amdgcn_sched_barrier(0);
    prefetch(((int64_t)job_id * g.n_ty3 + ty) * g.n_tx3 + tx_begin);

    for (int tx = tx_begin; tx < tx_end; ++tx) {
        const int tileX = g.avs - kXT / 2 + (kXT / 2) * tx;

        const int64_t tile = ((int64_t)job_id * g.n_ty3 + ty) * g.n_tx3 + tx;
        CHD_PHASE(9);
        if constexpr (kForward) {
            // ---- forward x: two rows per lane -> x bins 2..6 ----------------
            float l1 = 0.0f;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float s[16];
                unpack16(rows[h][0], rows[h][1], g.black, s);
                c32 X[5];
                l1 += xdft16(s, X) * roww[h];
                const int m = (lane + 64 * h) & 15;
#pragma unroll
                for (int j = 0; j < 5; ++j) L.u.A[rz[h] * 5 + j][m] = c32{X[j].x * roww[h], X[j].y * roww[h]};
            }
            // the same in every lane: a scalar, so that the mask steps branch
            // on it uniformly
            const float erru = kErrK * kU * wave_total(l1);
            __syncthreads();
            CHD_PHASE(0);

            // ---- vertical: 16-point DFT of the data lines + parity twiddle ---
            // (with alternating line parity the twiddle of the odd-parity
            // planes is applied in the temporal stage instead: there a lane
            // holds one ky, and only 4 of its 8 values need it)
            {
                c32 v[16];
                const bool act = lane < 40;
                if (act) {
#pragma unroll
                    for (int m = 0; m < 16; ++m) v[m] = L.u.A[lane][m];
                }
                __syncthreads();  // A and S share storage
                if (act) {
                    const int z = lane / 5, k = lane - 5 * (lane / 5);
                    fft<16, -1>(v);
                    if (alt) {
#pragma unroll
                        for (int ky = 0; ky < 16; ++ky) Sat(L, k, ky, z) = v[ky];
                    } else {
                        const bool odd = fy_odd;
#pragma unroll
                        for (int ky = 0; ky < 16; ++ky) {
                            c32 x0 = v[ky];
                            if (odd) x0 = twiddle<-1>(x0, ky);
                            Sat(L, k, ky, z) = x0;
                            Sat(L, k, ky + 16, z) = odd ? c32{-x0.x, -x0.y} : x0;
                        }
                    }
                }
            }
            __syncthreads();
            CHD_PHASE(1);

            // ---- temporal: 8-point DFT per (k, ky) ----------------------------
            if (alt) {
                for (int c = lane; c < 80; c += 64) {
                    const int k = c >> 4, ky = c & 15;
                    c32 v[8];
#pragma unroll
                    for (int z = 0; z < 8; ++z) v[z] = Sat(L, k, ky, z);
                    // parity twiddle exp(-2 pi i ky / 32) of the odd-parity planes
                    // (odd z when plane 0 has even parity, even z otherwise)
                    auto turn = [&](c32 b) {
                        return c32{b.x * ptw.x - b.y * ptw.y, b.x * ptw.y + b.y * ptw.x};
                    };
                    if (pbits & 1) {
                        asm volatile("");   // a scalar branch, not selects
#pragma unroll
                        for (int z = 0; z < 8; z += 2) v[z] = turn(v[z]);
                    } else {
                        asm volatile("");
#pragma unroll
                        for (int z = 1; z < 8; z += 2) v[z] = turn(v[z]);
                    }
                    fft<8, -1>(v);
                    if (half) {
#pragma unroll
                        for (int z = 0; z < 8; ++z) Sat(L, k, ky, z) = v[z];
                    } else {
#pragma unroll
                        for (int z = 0; z < 8; ++z) {
                            Sat(L, k, ky, z) = v[z];
                            Sat(L, k, ky + 16, z) = c32{sgn * v[(z + 4) & 7].x, sgn * v[(z + 4) & 7].y};
                        }
                    }
                }
            } else {
                for (int c = lane; c < 160; c += 64) {
                    const int k = c >> 5, ky = c & 31;
                    c32 v[8];
#pragma unroll
                    for (int z = 0; z < 8; ++z) v[z] = Sat(L, k, ky, z);
                    fft<8, -1>(v);
#pragma unroll
                    for (int z = 0; z < 8; ++z) Sat(L, k, ky, z) = v[z];
                }
            }
            __syncthreads();
            CHD_PHASE(2);

            if constexpr (MODE == kTileSpectra) {
                load_rows(min(tx + 1, tx_end - 1));
                for (int i = lane; i < 1280; i += 64) {
                    const c32 v = Sat(L, i >> 8, (i >> 3) & 31, i & 7);
                    spec[tile * 1280 + i] = make_float2(v.x, v.y);
                }
                __syncthreads();
                continue;
            }
            if constexpr (MODE == kTileFeatures) {
                load_rows(min(tx + 1, tx_end - 1));
                float *feat = reinterpret_cast<float *>(spec) + tile * 640;
                for (int i = lane; i < 640; i += 64) {
                    const c32 v = Sat(L, i >> 7, (i >> 3) & 15, i & 7);
                    feat[i] = chd::mag_feature(v.x, v.y);
                }
                __syncthreads();
                continue;
            }
            if constexpr (MODE == kTileApplyFwd) {
                // the tile's soft mask (loaded during the previous tile)
#pragma unroll
                for (int j = 0; j < kMaskPerLane; ++j) L.m[lane + 64 * j] = pre_m[j];
                __syncthreads();
                // the next tile's rows, then its mask (the current tile's
                // again on the last tile: the loads sit on every path)
                load_rows(min(tx + 1, tx_end - 1));
                __builtin_amdgcn_sched_barrier(0);   // the rows first: the next header waits for them only
                prefetch(tile + (tx + 1 < tx_end ? 1 : 0));
            } else {

                // ---- carrier-mirror mask (transform_pal_3d.cpp:277-330) -----------
                // visit v = (z*32 + y)*3 + (x-2) has threshold thr[v].  Decisions are
                // gathered by wave ballots into keep words (plane (x-2, z), bit y);
                // a kept visit also keeps its partner, whose word is the mirror of
                // the ballot.  Keep is the OR over visits, as in the reference.
                {
                    // Step i (unrolled) decides plane (x, z) = (2 + i/4, i%4) on
                    // lanes 0-31 and plane (x, z + 4) on lanes 32-63, so the
                    // half enters every offset as a lane term and the rest is an
                    // immediate.  The mirror of plane z lies at z' = (10 - z) % 8,
                    // which is 4 above the lower half's for z < 3 and 4 below for
                    // z = 3.  The next step's bins are read while this one decides.
                    // The uniform-threshold case (the default) is a loop of its
                    // own with no memory access but LDS: a load under a branch in
                    // every step would put a wait for all outstanding memory
                    // operations after each join.
                    const int yo = vy * kSpad + 4 * vhalf;
                    const int yrp = ((40 - vy) & 31) * kSpad + 4 * vhalf, yrm = yrp - 8 * vhalf;
                    const int tho = vy * 3 + 384 * vhalf;
                    auto off_a = [&](int i) { return (i >> 2) * kSk + (i & 3) + yo; };
                    auto off_b = [&](int i) {
                        const int zr = (10 - (i & 3)) & 7;
                        return (4 - (i >> 2)) * kSk + zr + ((i & 3) < 3 ? yrp : yrm);
                    };
                    auto steps = [&](auto uc) {
                        constexpr bool kUniform = decltype(uc)::value;
                        // With one threshold for every visit, plane (x, z) = (4, 2)
                        // holds the pairs of plane (4, 0) with a and b swapped
                        // (z' = (10 - z) % 8, rows mirrored), and (4, 6) those of
                        // (4, 4).  decide32 is symmetric in a and b, so step 10 is
                        // step 8's ballot moved to the mirror rows, and its pairs
                        // are re-decided (if at all) through plane (4, 0) / (4, 4).
                        constexpr int kSteps = kUniform ? 11 : 12;
                        const DecideConst qu = decide_const(erru, g.thr_uniform);
                        auto step_of = [](int j) { return (kUniform && j >= 10) ? j + 1 : j; };
                        c32 a = L.u.S[off_a(0)], b = L.u.S[off_b(0)];
#pragma unroll
                        for (int j = 0; j < kSteps; ++j) {
                            const int i = step_of(j);
                            c32 an = a, bn = b;
                            if (j + 1 < kSteps) {
                                an = L.u.S[off_a(step_of(j + 1))];
                                bn = L.u.S[off_b(step_of(j + 1))];
                            }
                            const int k = i >> 2, z = i & 3;
                            const DecideConst q = kUniform ? qu : decide_const(erru, thr32[z * 96 + k + tho]);
                            uint64_t kb, fb;
                            decide32_ballots(a, b, q, kb, fb);
                            if (lane == 0) {
                                // planes (k, z) on lanes 0-31 and (k, z + 4) on lanes 32-63
                                *reinterpret_cast<uint64_t *>(&L.bal[pw(k, z)]) = kb;
                                *reinterpret_cast<uint64_t *>(&L.flg[pw(k, z)]) = fb;
                                if (kUniform && i == 8) {
                                    const uint64_t m = mirror_rows((uint32_t)kb) |
                                                       ((uint64_t)mirror_rows((uint32_t)(kb >> 32)) << 32);
                                    *reinterpret_cast<uint64_t *>(&L.bal[pw(k, 2)]) = m;
                                    *reinterpret_cast<uint64_t *>(&L.flg[pw(k, 2)]) = 0u;
                                }
                            }
                            a = an;
                            b = bn;
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    };
                    // The half form: planes z < 4 only, rows ky >= 16 read from
                    // their stored images, rows ky - 16 at kz + 4 (bins of equal
                    // magnitude).  Steps 0-3 decide (x, z) = (2, s) on lanes 0-31
                    // and (3, s) on lanes 32-63.  Step 4 decides the x = 4 pairs,
                    // each once: plane (4, 0) on lanes 0-31 (plane (4, 2) holds
                    // its pairs, a and b swapped, as in the full form); plane
                    // (4, 1), whose visit y and 8 - y share a pair, on lanes 32-46
                    // for y = 5..19 (its self-mirror bins y = 4, 20 are kept
                    // anyway); plane (4, 3), whose visit y and 24 - y share a pair
                    // through the image rows, on lanes 47-63 for y = 12..28.  The
                    // other visits of those planes take the decision of their
                    // pair's (a mirrored ballot); only the deciding visit is
                    // flagged, and its FP64 re-decision sets both bins.
                    auto steps_half = [&]() {
                        const int yr = (40 - vy) & 31;
                        const int a0 = (vy & 15) * kSpad + 4 * (vy >> 4);
                        // mirror z' = (10 - z) % 8 is 2 - z for z < 3 and 7 for
                        // z = 3; a mirror row yr >= 16 adds 4 (mod 8) to it
                        const int b0 = (yr & 15) * kSpad + 4 * (yr >> 4), b0m = b0 - 8 * (yr >> 4);
                        // step 4: this lane's visit (z4, y4) of x = 4 and its mirror
                        const int y4 = lane < 32 ? lane : lane < 47 ? lane - 27 : lane - 35;
                        const int z4 = lane < 32 ? 0 : lane < 47 ? 1 : 3;
                        const int yr4 = (40 - y4) & 31, zr4 = (10 - z4) & 7;
                        const int a4 = 2 * kSk + (y4 & 15) * kSpad + z4 + 4 * (y4 >> 4);
                        const int b4 = 2 * kSk + (yr4 & 15) * kSpad + ((zr4 + 4 * (yr4 >> 4)) & 7);
                        auto off_a = [&](int st) { return st < 4 ? a0 + vhalf * kSk + st : a4; };
                        auto off_b = [&](int st) {
                            return st < 3    ? b0 - vhalf * kSk + 4 * kSk + (2 - st)
                                   : st == 3 ? b0m - vhalf * kSk + 4 * kSk + 7
                                             : b4;
                        };
                        const DecideConst q = decide_const(erru, g.thr_uniform);
                        c32 a = L.u.S[off_a(0)], b = L.u.S[off_b(0)];
#pragma unroll
                        for (int st = 0; st < 5; ++st) {
                            c32 an = a, bn = b;
                            if (st + 1 < 5) {
                                an = L.u.S[off_a(st + 1)];
                                bn = L.u.S[off_b(st + 1)];
                            }
                            uint64_t kb, fb;
                            decide32_ballots(a, b, q, kb, fb);
                            if (lane == 0) {
                                if (st < 4) {
                                    L.bal[pw(0, st)] = (uint32_t)kb;
                                    L.flg[pw(0, st)] = (uint32_t)fb;
                                    L.bal[pw(1, st)] = (uint32_t)(kb >> 32);
                                    L.flg[pw(1, st)] = (uint32_t)(fb >> 32);
                                } else {
                                    auto rot16 = [](uint32_t w) { return (w << 16) | (w >> 16); };
                                    const uint32_t w1 = (uint32_t)((kb >> 32) & 0x7FFFu) << 5;
                                    const uint32_t w3 = (uint32_t)((kb >> 47) & 0x1FFFFu) << 12;
                                    L.bal[pw(2, 0)] = (uint32_t)kb;
                                    L.flg[pw(2, 0)] = (uint32_t)fb;
                                    L.bal[pw(2, 2)] = mirror_rows((uint32_t)kb);
                                    L.flg[pw(2, 2)] = 0u;
                                    L.bal[pw(2, 1)] = w1 | mirror_rows(w1);
                                    L.flg[pw(2, 1)] = (uint32_t)((fb >> 32) & 0x7FFFu) << 5;
                                    L.bal[pw(2, 3)] = w3 | rot16(mirror_rows(w3));
                                    L.flg[pw(2, 3)] = (uint32_t)((fb >> 47) & 0x1FFFFu) << 12;
                                }
                            }
                            a = an;
                            b = bn;
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    };
                    if (erru == 0.0f) {
                        // every windowed sample is zero, so is every bin: all pairs
                        // are kept (0 < 0 is false), none needs re-deciding
                        if (lane < 24) {
                            L.bal[lane] = ~0u;
                            L.flg[lane] = 0u;
                        }
                    } else if (half) {
                        steps_half();
                    } else if (g.thr_uniform >= 0.0f) {
                        steps(std::true_type{});
                    } else {
                        steps(std::false_type{});
                    }
                }
                __syncthreads();
                CHD_PHASE(3);
                // keep word (k, z) = own visits' ballot | ballot of the visit plane
                // whose partners lie on (k, z), moved to the mirror rows
                // (x_v = 8 - x, z_v = (10 - z) % 8); plus the self-mirror bins
                // (x = 4, y in {4, 20}, z in {1, 5}), which are always kept.
                constexpr uint32_t kSelf = (1u << 4) | (1u << 20);
                auto rot16 = [](uint32_t w) { return (w << 16) | (w >> 16); };
                if (lane < 40) {
                    const int k = lane >> 3, z = lane & 7;
                    uint32_t w;
                    if (half) {
                        // from the planes z < 4; plane (x, 7) is plane (x, 3)
                        // rotated, and word (x, z + 4) is word (x, z) rotated
                        const int zl = z & 3;
                        w = (k <= 2) ? L.bal[pw(k, zl)] : 0u;
                        if (k >= 2) w |= mirror_rows(zl < 3 ? L.bal[pw(4 - k, 2 - zl)] : rot16(L.bal[pw(4 - k, 3)]));
                        if (k == 2 && zl == 1) w |= kSelf;
                        if (z >= 4) w = rot16(w);
                    } else {
                        w = (k <= 2) ? L.bal[pw(k, z)] : 0u;
                        if (k >= 2) w |= mirror_rows(L.bal[pw(4 - k, (10 - z) & 7)]);
                        if (k == 2 && (z == 1 || z == 5)) w |= kSelf;
                    }
                    L.keep[lane] = w;
                }
                // flag words (x - 2) * 8 + z on lanes 0-23, self-mirror bins
                // removed (the half form flags planes z < 4 only)
                const bool flag_lane = lane < 24 && (!half || (lane & 7) < 4);
                const uint32_t myflag =
                    flag_lane ? (L.flg[pw(lane >> 3, lane & 7)] & ((lane == 17 || lane == 21) ? ~kSelf : ~0u)) : 0u;
                const uint64_t fwords = __ballot(myflag != 0u);
                __syncthreads();
                CHD_PHASE(4);
                CHD_PHASE_COUNT(16, fwords != 0u);
                CHD_PHASE_COUNT(17, __popcll(fwords));

                // ---- FP64 re-decision of comparisons FP32 cannot settle ----------
                // Each lane forms the FP64 DFT of its two rows at bins x and 8 - x
                // once per flagged x; a flagged visit turns them by the (ky, kz)
                // phase of its bins and a wave reduction gives the exact bins.  x is
                // a compile-time constant, so the row DFT's twiddles are immediates.
                if (fwords != 0u) {
                    const double black = (double)g.black;
                    auto bins = [&](auto xc) {
                        constexpr int X = decltype(xc)::value, XR = 8 - X;
                        uint32_t zmask = (uint32_t)(fwords >> (8 * (X - 2))) & 0xFFu;
                        if (zmask == 0u) return;
                        // With c, s the cos/sin of 2 pi X n / 16, bin X of a row is
                        // sum u (c - i s) and bin 8-X is sum (-1)^n u (c + i s): split
                        // by n parity.  A row outside the tile's lines holds black:
                        // u = 0.
                        // one row at a time: both rows' FP64 samples at once would
                        // not fit the tile kernel's register budget
                        double ra[2], ia[2], rb[2], ib[2];
#pragma unroll 1
                        for (int h = 0; h < 2; ++h) {
                            const uint4 q0 = h ? rows[1][0] : rows[0][0], q1 = h ? rows[1][1] : rows[0][1];
                            const uint32_t raw[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
                            // the row's z and y window factor, applied to the sums
                            const double rw = (h ? rowp[1] : rowp[0]) != nullptr
                                                  ? L.wz[h ? rz[1] : rz[0]] * L.wy[h ? ry[1] : ry[0]]
                                                  : 0.0;
                            // opaque per bin pair: otherwise the 32 windowed samples
                            // are shared by the three pairs and held in registers
                            double blk = black;
                            asm volatile("" : "+v"(blk));
                            double ce = 0.0, co = 0.0, se = 0.0, so = 0.0;
#pragma unroll
                            for (int n = 0; n < 16; n += 2) {
                                const uint32_t w2 = raw[n >> 1];
                                const double u0 = ((double)(w2 & 0xFFFFu) - blk) * c_tab.wx_d[n];
                                const double u1 = ((double)(w2 >> 16) - blk) * c_tab.wx_d[n + 1];
                                ce = fma(u0, kCos16[(X * n) & 15], ce);
                                se = fma(u0, kSin16[(X * n) & 15], se);
                                co = fma(u1, kCos16[(X * (n + 1)) & 15], co);
                                so = fma(u1, kSin16[(X * (n + 1)) & 15], so);
                            }
                            const double r0 = (ce + co) * rw, i0 = -(se + so) * rw, r1 = (ce - co) * rw,
                                         i1 = (se - so) * rw;
                            if (h == 0) {
                                ra[0] = r0; ia[0] = i0; rb[0] = r1; ib[0] = i1;
                            } else {
                                ra[1] = r0; ia[1] = i0; rb[1] = r1; ib[1] = i1;
                            }
                        }
                        while (zmask != 0u) {
                            const int z = __builtin_ctz(zmask);
                            zmask &= zmask - 1u;
                            uint32_t bits = (uint32_t)__builtin_amdgcn_readlane((int)myflag, (X - 2) * 8 + z);
                            const int zr = (10 - z) & 7;
                            while (bits != 0u) {
                                const int y = __builtin_ctz(bits);
                                bits &= bits - 1u;
                                const int yr = (40 - y) & 31;
                                double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
#pragma unroll
                                for (int h = 0; h < 2; ++h) {
                                    // turn by the (ky, kz) phase of the row: 4 kz z + ky y (32ths)
                                    const int pa = (4 * z * rz[h] + y * ry[h]) & 31;
                                    const int pb = (4 * zr * rz[h] + yr * ry[h]) & 31;
                                    const double ca = L.cs[pa], sa = L.sn[pa], cb = L.cs[pb], sb = L.sn[pb];
                                    a0 += ra[h] * ca + ia[h] * sa;
                                    a1 += ia[h] * ca - ra[h] * sa;
                                    b0 += rb[h] * cb + ib[h] * sb;
                                    b1 += ib[h] * cb - rb[h] * sb;
                                }
                                wave_sum4_to63(a0, a1, b0, b1);
                                const double th =
                                    g.thr_uniform >= 0.0f ? g.thr_uniform_d : thr64[(z * 32 + y) * 3 + (X - 2)];
                                if (lane == 63 && decide64(a0, a1, b0, b1, th)) {
                                    L.keep[(X - 2) * 8 + z] |= 1u << y;
                                    L.keep[(XR - 2) * 8 + zr] |= 1u << yr;
                                    if (half) {
                                        // and the image visit (z + 4, y + 16)
                                        L.keep[(X - 2) * 8 + z + 4] |= 1u << ((y + 16) & 31);
                                        L.keep[(XR - 2) * 8 + ((zr + 4) & 7)] |= 1u << ((yr + 16) & 31);
                                    }
                                }
                            }
                        }
                    };
                    bins(std::integral_constant<int, 2>{});
                    bins(std::integral_constant<int, 3>{});
                    bins(std::integral_constant<int, 4>{});
                    __syncthreads();
                }
                CHD_PHASE(5);
                // the next tile's rows, in flight through the inverse (the
                // current tile's again on the last tile: the loads sit on every
                // path)
                load_rows(min(tx + 1, tx_end - 1));
            }  // hard mask
        } else {
#pragma unroll
            for (int j = 0; j < kSpecPerLane; ++j) {
                const int i = lane + 64 * j;
                Sat(L, i >> 8, (i >> 3) & 31, i & 7) = c32{pre_s[j].x, pre_s[j].y};
            }
#pragma unroll
            for (int j = 0; j < kMaskPerLane; ++j) L.m[lane + 64 * j] = pre_m[j];
            __syncthreads();
            prefetch(tile + (tx + 1 < tx_end ? 1 : 0));
        }

        // ---- masked inverse z: outputs z = 0,1,4,5 ------------------------
        // (rows ky < 16 only in the half form, on two passes; there the lane's
        // ky is lane % 16 and its turn exp(2 pi i ky / 32) is conj(ptw))
        if (half) {
            const c32 htw = {ptw.x, -ptw.y};
            for (int c = lane; c < 80; c += 64) {
                const int k = c >> 4, ky = c & 15;
                // the paired masks' visit of bin (k, ky, z): its own visit
                // (x <= 4) or the one it is the partner of, moved to planes
                // z < 4 (visit (z + 4, y + 16) is visit (z, y)); with z a
                // constant, a lane base plus an immediate
                const int mown = ky * 3 + k, mlo = ((40 - ky) & 31) * 3 + (4 - k), mhi = (24 - ky) * 3 + (4 - k);
                c32 v[8];
#pragma unroll
                for (int z = 0; z < 8; ++z) {
                    const c32 t = Sat(L, k, ky, z);
                    if constexpr (MODE == kTileApplyFwd) {
                        const int zm = (10 - z) & 7;
                        const int own = mown + (z & 3) * 96 + (z >= 4 ? 48 : 0);
                        const int mir = zm >= 4 ? mhi + (zm - 4) * 96 : mlo + zm * 96;
                        const float mv = L.m[k <= 2 ? own : mir];
                        v[z] = c32{t.x * mv, t.y * mv};
                    } else {
                        v[z] = ((L.keep[k * 8 + z] >> ky) & 1u) ? t : c32{0.0f, 0.0f};
                    }
                }
                fft<8, +1>(v);
                auto turn = [&](c32 b) { return c32{b.x * htw.x - b.y * htw.y, b.x * htw.y + b.y * htw.x}; };
                if (pbits & 1) {
                    asm volatile("");
                    v[0] = turn(v[0]);
                    v[4] = turn(v[4]);
                } else {
                    asm volatile("");
                    v[1] = turn(v[1]);
                    v[5] = turn(v[5]);
                }
                Sat(L, k, ky, 0) = v[0];
                Sat(L, k, ky, 1) = v[1];
                Sat(L, k, ky, 2) = v[4];
                Sat(L, k, ky, 3) = v[5];
            }
        } else
        for (int c = lane; c < 160; c += 64) {
            const int k = c >> 5, ky = c & 31;
            c32 v[8];
#pragma unroll
            for (int z = 0; z < 8; ++z) {
                const c32 t = Sat(L, k, ky, z);
                if constexpr (kSoft) {
                    const float mv = L.m[mask_visit(k, ky, z)];
                    v[z] = c32{t.x * mv, t.y * mv};
                } else {
                    v[z] = ((L.keep[k * 8 + z] >> ky) & 1u) ? t : c32{0.0f, 0.0f};
                }
            }
            fft<8, +1>(v);
            // Output planes of odd line parity are turned by exp(2 pi i ky / 32)
            // here, where the lane holds one ky: the inverse-y fold then adds
            // rows ky and ky + 16 for every plane (the turn of ky + 16 is minus
            // that of ky).
            auto turn = [&](c32 b) { return c32{b.x * itw.x - b.y * itw.y, b.x * itw.y + b.y * itw.x}; };
            // (scalar branches; the empty asm keeps the compiler from turning
            // them into selects over all four planes)
            if (alt) {
                if (pbits & 1) {
                    asm volatile("");
                    v[0] = turn(v[0]);
                    v[4] = turn(v[4]);
                } else {
                    asm volatile("");
                    v[1] = turn(v[1]);
                    v[5] = turn(v[5]);
                }
            } else {
                if (pbits & 0x01) v[0] = turn(v[0]);
                if (pbits & 0x02) v[1] = turn(v[1]);
                if (pbits & 0x10) v[4] = turn(v[4]);
                if (pbits & 0x20) v[5] = turn(v[5]);
            }
            Sat(L, k, ky, 0) = v[0];
            Sat(L, k, ky, 1) = v[1];
            Sat(L, k, ky, 2) = v[4];
            Sat(L, k, ky, 3) = v[5];
        }
        __syncthreads();
        CHD_PHASE(6);

        // ---- inverse y: fold 32 -> 16 bins (rows ky, ky + 16 of the plane) ----
        // Two lanes per 16-point transform (plane zs, x bin k): lane 2p + h
        // transforms the inputs of parity h (8 points), the odd half is turned
        // by exp(2 pi i m / 16), and the pair swaps halves (a DPP move) to form
        // outputs 8h .. 8h + 7: the last radix-2 stage of fft<16>, the same
        // operations in the same order.
        {
            c32 v[8];
            const int pr = lane >> 1, h = lane & 1;
            const bool act = lane < 40;
            const int zs = pr / 5, k = pr - 5 * (pr / 5);
            if (act) {
                if (half) {
                    // the fold is 2 T[ky], the 2 folded into the x scale
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[j] = Sat(L, k, 2 * j + h, zs);
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[j] = cadd(Sat(L, k, 2 * j + h, zs), Sat(L, k, 2 * j + h + 16, zs));
                }
            }
            __syncthreads();  // A and S share storage
            if (act) {
                fft<8, +1>(v);
                if (h) {
#pragma unroll
                    for (int m = 1; m < 8; ++m) v[m] = twiddle<+1>(v[m], 2 * m);
                }
                // lane h = 0: E + W O (outputs 0..7); h = 1: E - W O (8..15)
                const float sg = h ? -1.0f : 1.0f;
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const float ox = swap_pair(v[m].x), oy = swap_pair(v[m].y);
                    L.u.A[pr][8 * h + m] = c32{fmaf(sg, v[m].x, ox), fmaf(sg, v[m].y, oy)};
                }
            }
        }
        __syncthreads();
        CHD_PHASE(7);

        // ---- inverse x (c2r on bins 2..6) and overlap-add ----------------
        {
            c32 C[5];
            // (plain mode: the scale is in the forward row weights)
            constexpr float sc = kFoldScale ? 1.0f : kInvScale;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const c32 t = L.u.A[zsel * 5 + k][om];
                C[k] = sc == 1.0f ? t : c32{t.x * sc, t.y * sc};
            }
            float out[16];
            xidft16(C, out);
            // Columns tileX .. tileX+7 are complete (tiles tx-1 and tx).
            // Every lane stores on every tile: a lane with nothing to write
            // (a row outside the tile's lines, the strip's leading tile)
            // stores to columns 0..31 of its plane's first row, which no
            // kernel reads.  With at least the same stores on every path,
            // the next tile's loads are waited for by count, not together
            // with these stores.
            // Staged (the paired learned path's masked inverse): tiles
            // 4j .. 4j+2 of a strip (its first owned tile is a multiple of 4,
            // its columns start a 128-byte line) go to LDS, and tile 4j+3, or
            // the strip's last tile, stores the group's columns, whole lines
            // of each row.  That path also streams the masks in; with both
            // the partial-line stores of the direct form and the mask loads,
            // the kernel ran 1.5x slower.  The plain kernel, which loads
            // only the composite, runs faster with the direct stores.
            const bool owned = tx >= tx_own0;
            const bool put = owned && orow_ok;
            {
                const int c0 = tileX - g.avs + kColOrigin;
                float *dst = put ? orow + c0 : orow_none;
                float v8[8];
                if (tileX >= g.avs && tileX + 8 <= g.ave) {
                    // an interior tile: every column is active (a scalar
                    // branch; the edge tiles take the per-column test)
#pragma unroll
                    for (int n = 0; n < 8; ++n) v8[n] = carry[n] + out[n];
                } else {
#pragma unroll
                    for (int n = 0; n < 8; ++n) {
                        const int col = tileX + n;
                        v8[n] = (col >= g.avs && col < g.ave) ? carry[n] + out[n] : 0.0f;
                    }
                }
                if constexpr (kStaged) {
                    const int slot = tx & 3;
                    if (owned && slot < 3 && tx + 1 < tx_end) {
                        // (a scalar branch) stage this tile's columns
                        L.stage[slot][0][lane] = make_float4(v8[0], v8[1], v8[2], v8[3]);
                        L.stage[slot][1][lane] = make_float4(v8[4], v8[5], v8[6], v8[7]);
                        gstore_f4(orow_none, v8[0], v8[1], v8[2], v8[3]);
                        gstore_f4(orow_none + 4, v8[4], v8[5], v8[6], v8[7]);
                    } else if (owned) {
                        // the group's staged tiles, then this one
                        float *gd = put ? dst - 8 * slot : orow_none;
                        for (int i = 0; i < slot; ++i) {
                            const float4 a = L.stage[i][0][lane], b = L.stage[i][1][lane];
                            gstore_f4(gd + 8 * i, a.x, a.y, a.z, a.w);
                            gstore_f4(gd + 8 * i + 4, b.x, b.y, b.z, b.w);
                        }
                        gstore_f4(gd + 8 * slot, v8[0], v8[1], v8[2], v8[3]);
                        gstore_f4(gd + 8 * slot + 4, v8[4], v8[5], v8[6], v8[7]);
                    } else {
                        gstore_f4(orow_none, v8[0], v8[1], v8[2], v8[3]);
                        gstore_f4(orow_none + 4, v8[4], v8[5], v8[6], v8[7]);
                    }
                } else {
                    gstore_f4(dst, v8[0], v8[1], v8[2], v8[3]);
                    gstore_f4(dst + 4, v8[4], v8[5], v8[6], v8[7]);
                }
                if (put && tx == g.n_tx3 - 1) {
#pragma unroll
                    for (int n = 0; n < 8; ++n) {
                        const int col = tileX + 8 + n;
                        v8[n] = (col >= g.avs && col < g.ave) ? out[8 + n] : 0.0f;
                    }
                    gstore_f4(dst + 8, v8[0], v8[1], v8[2], v8[3]);
                    gstore_f4(dst + 12, v8[4], v8[5], v8[6], v8[7]);
                }
            }
#pragma unroll
            for (int n = 0; n < 8; ++n) carry[n] = out[8 + n];
        }
        __syncthreads();
        CHD_PHASE(8);
        CHD_PHASE_COUNT(18, 1);
    }
    }  // work items
    CHD_PHASE_FLUSH();
    if constexpr (PERSIST) {
        if (lane0 == 0 && atomicAdd(&wq[8], 1u) == (unsigned)G - 1u) {
#pragma unroll
            for (int i = 0; i < 9; ++i) atomicExch(&wq[i], 0u);
        }
    }
}

// ---------------------------------------------------------------------------
// 2D tile kernel (TransformPal2D, transform_pal_2d.cpp:107-287).  One
// 64-lane workgroup = 4 groups of 16 lanes; group q runs its own strip of
// kStrip2 (+1 leading) tiles of one tile row of one field, one line per lane.
// ---------------------------------------------------------------------------

// Rows are padded to 17: in the vertical stages lane (q, k) walks row
// (q, k), and with 16-entry rows every active lane would hit the same banks.
constexpr int kA2pad = 17;
struct Lds2 {
    c32 A[4][9][kA2pad];      // per group: (k, line) after x DFT / before x inverse
    uint32_t kw[4][9];        // per group: keep word of bin k, bit = line
    uint16_t flag[4][80];
    int nflag[4];
};

// Bit y of each 16-bit group of w moved to bit (24 - y) % 16: the 2D mask's
// mirror row, for the four groups of a wave at once.
__device__ __forceinline__ uint64_t mirror16x4(uint64_t w)
{
    w = __builtin_bitreverse64(w);   // group q to 3 - q, bit y to 15 - y
    w = (w >> 32) | (w << 32);
    w = ((w >> 16) & 0x0000FFFF0000FFFFull) | ((w & 0x0000FFFF0000FFFFull) << 16);   // groups back in place
    return ((w << 9) & 0xFE00FE00FE00FE00ull) | ((w >> 7) & 0x01FF01FF01FF01FFull);   // 15 - y to (24 - y) % 16
}

__device__ __forceinline__ void load32(const uint16_t *p, float black, float (&s)[32])
{
    uint4 q[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) q[h] = gload_u4(p + 8 * h);
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const uint4 u = q[h];
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            s[8 * h + 2 * i] = (float)(w[i] & 0xFFFFu) - black;
            s[8 * h + 2 * i + 1] = (float)(w[i] >> 16) - black;
        }
    }
}

// Sums over each group of 16 lanes (one DPP row), in every lane of the
// group: pairs, quads, half rows, rows.
__device__ __forceinline__ float group_sum16(float v)
{
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    v += dpp_f<0x141>(v);
    v += dpp_f<0x140>(v);
    return v;
}
__device__ __forceinline__ void group_sum16_4d(double &a, double &b, double &c, double &d)
{
    a += dpp_d<0xB1>(a); b += dpp_d<0xB1>(b); c += dpp_d<0xB1>(c); d += dpp_d<0xB1>(d);
    a += dpp_d<0x4E>(a); b += dpp_d<0x4E>(b); c += dpp_d<0x4E>(c); d += dpp_d<0x4E>(d);
    a += dpp_d<0x141>(a); b += dpp_d<0x141>(b); c += dpp_d<0x141>(c); d += dpp_d<0x141>(d);
    a += dpp_d<0x140>(a); b += dpp_d<0x140>(b); c += dpp_d<0x140>(c); d += dpp_d<0x140>(d);
}

// Windowed 32-sample real DFT to bins 4..12 (X[j] = bin j+4), folded by the
// half-period symmetry: even bins see e[n] = u[n] + u[n+16], odd bins
// o[n] = u[n] - u[n+16] (n < 16).  The even bins fold again as xdft16's do
// (bins 4, 8, 12 on e[n] + e[n+8], bins 6, 10 on e[n] - e[n+8]), and bins k
// and 16 - k share their sums: at t = 2 pi k n / 32 the twiddle of 16 - k is
// (-1)^n times the conjugate of k's, so both come from the even-n and odd-n
// sums of v[n] cos t and -v[n] sin t.  Returns sum |u[n]|.
__device__ __forceinline__ float xdft32(const float (&s)[32], c32 (&X)[9])
{
    float u[32], l1 = 0.0f;
#pragma unroll
    for (int n = 0; n < 32; ++n) {
        u[n] = s[n] * c_tab.wx2[n];
        l1 += fabsf(u[n]);
    }
    float e[16], o[16];
#pragma unroll
    for (int n = 0; n < 16; ++n) {
        e[n] = u[n] + u[n + 16];
        o[n] = u[n] - u[n + 16];
    }
    float f[8], g[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        f[n] = e[n] + e[n + 8];
        g[n] = e[n] - e[n + 8];
    }
    // bin 8: sum f[n] (-i)^n
    X[4] = c32{(f[0] - f[2]) + (f[4] - f[6]), (f[3] - f[1]) + (f[7] - f[5])};
    // bins 4 and 12: p[n] = f[n] - f[n+4]
    const float p0 = f[0] - f[4], p1 = f[1] - f[5], p2 = f[2] - f[6], p3 = f[3] - f[7];
    const float r = 0.70710678118654752f;
    const float a = (p1 - p3) * r, b = (p1 + p3) * r;
    X[0] = c32{p0 + a, -(p2 + b)};
    X[8] = c32{p0 - a, p2 - b};
    // bins 6 and 10 on g[n] (8 terms), then 5 and 11, 7 and 9 on o[n] (16)
    auto pair = [](const float *v, int nv, int k, c32 &lo, c32 &hi) {
        float ce = 0.0f, co = 0.0f, se = 0.0f, so = 0.0f;
#pragma unroll
        for (int n = 0; n < nv; n += 2) {
            ce = cmac(ce, v[n], cos32(k * n));
            se = cmac(se, v[n], -sin32(k * n));
            co = cmac(co, v[n + 1], cos32(k * (n + 1)));
            so = cmac(so, v[n + 1], -sin32(k * (n + 1)));
        }
        lo = c32{ce + co, se + so};
        hi = c32{ce - co, so - se};
    };
    pair(g, 8, 6, X[2], X[6]);
    pair(o, 16, 5, X[1], X[7]);
    pair(o, 16, 7, X[3], X[5]);
    return l1;
}

// c2r over bins 4..12 (C[j] = bin j+4, pre-scaled): out[n] = 2 Re sum C e^{+i..}.
// Even bins repeat every 16 samples, odd bins change sign.
__device__ __forceinline__ void xidft32(const c32 (&C)[9], float (&out)[32])
{
    // As xidft16: samples n = 0 and 8 directly, n = 1..7 with 16 - n from
    // the even-k and odd-k sums of C.x cos t and C.y sin t (t = 2 pi k n / 32).
#pragma unroll
    for (int n = 0; n <= 8; n += 8) {
        float E = 0.0f, O = 0.0f;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int k = j + 4;
            const float c = cos32(k * n), sn = sin32(k * n);
            if (k & 1) {
                O = cmac(O, C[j].x, c);
                O = cmac(O, C[j].y, -sn);
            } else {
                E = cmac(E, C[j].x, c);
                E = cmac(E, C[j].y, -sn);
            }
        }
        out[n] = E + O;
        out[n + 16] = E - O;
    }
#pragma unroll
    for (int n = 1; n < 8; ++n) {
        float pc = 0.0f, ps = 0.0f, qc = 0.0f, qs = 0.0f;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int k = j + 4;
            const float c = cos32(k * n), sn = sin32(k * n);
            if (k & 1) {
                qc = cmac(qc, C[j].x, c);
                qs = cmac(qs, C[j].y, sn);
            } else {
                pc = cmac(pc, C[j].x, c);
                ps = cmac(ps, C[j].y, sn);
            }
        }
        const float En = pc - ps, Em = pc + ps;       // E at n and 16 - n
        const float On = qc - qs, Om = -(qc + qs);    // O at n and 16 - n
        out[n] = En + On;
        out[n + 16] = En - On;
        out[16 - n] = Em + Om;
        out[32 - n] = Em - Om;
    }
}

__global__ void __launch_bounds__(64)
tpal2d_tiles(const Geom g, const FieldJob2D *__restrict__ jobs, int n_ty_max,
             const float *__restrict__ thr32, const double *__restrict__ thr64)
{
    __shared__ Lds2 L;
    const int lane = threadIdx.x;
    const int q = lane >> 4, ln = lane & 15;
    const int groups_per_row = (g.n_strips2 + 3) / 4;
    const int per_field = n_ty_max * groups_per_row;
    // contiguous share of the work per XCD (workgroups go round-robin over
    // the 8 XCDs), so neighbouring tile rows of a field share one L2
    const int total = gridDim.x;
    const int xcd = blockIdx.x & 7, nq = total >> 3, nr = total & 7;
    const int w = xcd * nq + min(xcd, nr) + (blockIdx.x >> 3);
    const int job_id = w / per_field;
    const int rem = w - job_id * per_field;
    const int ty = rem / groups_per_row;
    const int strip = (rem - ty * groups_per_row) * 4 + q;
    const FieldJob2D &job = jobs[job_id];

    // SourceField::getFirstActiveLine / getLastActiveLine (source_field.h:86-91)
    const int firstFL = (g.first + 1 - job.offset) / 2;
    const int lastFL = ((g.last - job.offset) / 2) + 1;
    const int n_ty = (lastFL - (firstFL - kYT2 / 2) + kYT2 / 2 - 1) / (kYT2 / 2);
    const bool row_live = ty < n_ty;
    const int tileY = firstFL - kYT2 / 2 + (kYT2 / 2) * ty;
    const int startY = max(firstFL - tileY, 0);
    const int endY = min(lastFL - tileY, kYT2);

    const int tx_own0 = strip * kStrip2;
    const bool strip_live = row_live && strip < g.n_strips2;
    const int tx_end = strip_live ? min(tx_own0 + kStrip2, g.n_tx2) : 0;
    const int tx_begin = strip_live ? max(tx_own0 - 1, 0) : 0;
    // all groups iterate the same number of steps (the longest strip)
    const int steps = kStrip2 + 1;

    const bool row_ok = (ln >= startY) && (ln < endY);
    float *orow = job.partial + (int64_t)(ty & 1) * g.part_plane + (int64_t)(tileY + ln) * g.cpitch;
    float carry[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) carry[i] = 0.0f;

    for (int step = 0; step < steps; ++step) {
        const int tx = tx_begin + step;
        const bool live = strip_live && tx < tx_end;
        const int tileX = g.avs - kXT2 / 2 + (kXT2 / 2) * tx;

        // ---- forward x: one line per lane -> bins 4..12 -------------------
        float s[32];
        const bool have = live && job.field != nullptr && row_ok;
        if (have) {
            load32(job.field + (int64_t)(tileY + ln) * g.pitch + (tileX + g.lmargin), g.black, s);
        } else {
#pragma unroll
            for (int n = 0; n < 32; ++n) s[n] = 0.0f;
        }
        const float rw = c_tab.wy2[ln];
        c32 X[9];
        const float l1 = xdft32(s, X);
        const float err = kErrK2 * kU * group_sum16(l1 * rw);
#pragma unroll
        for (int k = 0; k < 9; ++k) L.A[q][k][ln] = c32{X[k].x * rw, X[k].y * rw};
        __syncthreads();

        // ---- vertical 16-point DFT: lanes 0..8 of each group ---------------
        if (ln < 9) {
            c32 v[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) v[m] = L.A[q][ln][m];
            fft<16, -1>(v);
#pragma unroll
            for (int m = 0; m < 16; ++m) L.A[q][ln][m] = v[m];
        }
        __syncthreads();

        // ---- mask: 80 visits per group (transform_pal_2d.cpp applyFilter) --
        // Step s decides visit (x, y) = (4 + s, ln) of every group against its
        // partner (16 - x, (24 - y) % 16), so the wave's keep ballot holds the
        // four groups' keep words of bin x (bit y), and moved to the mirror
        // rows those of bin 16 - x: a kept visit keeps both bins.  The
        // self-mirror bins (x = 8, y = 4, 12) are kept; a tile whose windowed
        // samples are all zero keeps every pair.  Unsure comparisons go to
        // the group's FP64 list (visit y * 5 + x - 4) in lane order.
        {
            const uint64_t zero = __builtin_amdgcn_ballot_w64(err == 0.0f);
            constexpr uint64_t kSelf2 = 0x1010101010101010ull;
            const DecideConst qu = decide_const(err, g.thr_uniform);
            const int yr = (24 - ln) & 15;
            uint64_t K[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) K[k] = 0;
            int nfl = 0;
#pragma unroll
            for (int st = 0; st < 5; ++st) {
                const int x = 4 + st, xr = 16 - x;
                const c32 a = L.A[q][x - 4][ln], b = L.A[q][xr - 4][yr];
                const DecideConst dq = g.thr_uniform >= 0.0f ? qu : decide_const(err, thr32[ln * 5 + st]);
                uint64_t kb, fb;
                decide32_ballots(a, b, dq, kb, fb);
                kb |= zero;
                fb &= ~zero;
                if (x == 8) {
                    kb |= kSelf2;
                    fb &= ~kSelf2;
                }
                K[x - 4] |= kb;
                K[xr - 4] |= mirror16x4(kb);
                const uint32_t fq = (uint32_t)(fb >> (16 * q)) & 0xFFFFu;
                if ((fq >> ln) & 1u) L.flag[q][nfl + __popc(fq & ((1u << ln) - 1u))] = (uint16_t)(ln * 5 + st);
                nfl += __popc(fq);
            }
            if (ln == 0) {
                L.nflag[q] = nfl;
#pragma unroll
                for (int k = 0; k < 9; ++k) L.kw[q][k] = (uint32_t)(K[k] >> (16 * q)) & 0xFFFFu;
            }
        }
        __syncthreads();

        // FP64 re-decision; the loop count is wave-uniform (max over groups).
        int nmax = 0;
#pragma unroll
        for (int gq = 0; gq < 4; ++gq) nmax = max(nmax, L.nflag[gq]);
        for (int f = 0; f < nmax; ++f) {
            const bool mine = f < L.nflag[q];
            const int v = mine ? L.flag[q][f] : 0;
            const int y = v / 5, x = 4 + (v - 5 * y);
            const int yr = (24 - y) & 15, xr = 16 - x;
            double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
            if (mine && have) {
                // this line's FP64 DFT at bins x and xr by Horner's rule in
                // w = e^{-2 pi i x / 32}, then turned by the line's vertical
                // phase e^{-2 pi i ky line / 16} (32ths: 2 ky line)
                const double rwd = c_tab.wy2_d[ln];
                const double wa0 = c_tab.cos32_d[x], wa1 = -c_tab.sin32_d[x];
                const double wb0 = c_tab.cos32_d[xr], wb1 = -c_tab.sin32_d[xr];
                double pa0 = 0, pa1 = 0, pb0 = 0, pb1 = 0;
#pragma unroll
                for (int n = 31; n >= 0; --n) {
                    const double xv = (double)s[n] * rwd * c_tab.wx2_d[n];
                    const double ta = pa0 * wa0 - pa1 * wa1 + xv, tb = pb0 * wb0 - pb1 * wb1 + xv;
                    pa1 = pa0 * wa1 + pa1 * wa0;
                    pb1 = pb0 * wb1 + pb1 * wb0;
                    pa0 = ta;
                    pb0 = tb;
                }
                const int ja = (2 * y * ln) & 31, jb = (2 * yr * ln) & 31;
                const double ca = c_tab.cos32_d[ja], sa = c_tab.sin32_d[ja];
                const double cb = c_tab.cos32_d[jb], sb = c_tab.sin32_d[jb];
                a0 = pa0 * ca + pa1 * sa;
                a1 = pa1 * ca - pa0 * sa;
                b0 = pb0 * cb + pb1 * sb;
                b1 = pb1 * cb - pb0 * sb;
            }
            group_sum16_4d(a0, a1, b0, b1);
            if (mine && ln == 0 && decide64(a0, a1, b0, b1, thr64[v])) {
                L.kw[q][x - 4] |= 1u << y;
                L.kw[q][xr - 4] |= 1u << yr;
            }
        }
        __syncthreads();

        // ---- masked inverse vertical ----------------------------------------
        if (ln < 9) {
            c32 v[16];
            const uint32_t kwl = L.kw[q][ln];
#pragma unroll
            for (int m = 0; m < 16; ++m) v[m] = ((kwl >> m) & 1u) ? L.A[q][ln][m] : c32{0.0f, 0.0f};
            fft<16, +1>(v);
#pragma unroll
            for (int m = 0; m < 16; ++m) L.A[q][ln][m] = v[m];
        }
        __syncthreads();

        // ---- inverse x and overlap-add ----------------------------------------
        c32 C[9];
        const float sc = 2.0f / 512.0f;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const c32 t = L.A[q][k][ln];
            C[k] = c32{t.x * sc, t.y * sc};
        }
        float out[32];
        xidft32(C, out);
        if (live && tx >= tx_own0 && row_ok) {
            const int c0 = tileX - g.avs + kColOrigin;
            float *dst = orow + c0;
            float v16[16];
            // interior tiles (all 16 columns active) skip the per-column test
            if (tileX >= g.avs && tileX + 16 <= g.ave) {
#pragma unroll
                for (int n = 0; n < 16; ++n) v16[n] = carry[n] + out[n];
            } else {
#pragma unroll
                for (int n = 0; n < 16; ++n) {
                    const int col = tileX + n;
                    v16[n] = (col >= g.avs && col < g.ave) ? carry[n] + out[n] : 0.0f;
                }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) gstore_f4(dst + 4 * i, v16[4 * i], v16[4 * i + 1], v16[4 * i + 2], v16[4 * i + 3]);
            if (tx == g.n_tx2 - 1) {
#pragma unroll
                for (int n = 0; n < 16; ++n) {
                    const int col = tileX + 16 + n;
                    v16[n] = (col >= g.avs && col < g.ave) ? out[16 + n] : 0.0f;
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    gstore_f4(dst + 16 + 4 * i, v16[4 * i], v16[4 * i + 1], v16[4 * i + 2], v16[4 * i + 3]);
            }
        }
#pragma unroll
        for (int n = 0; n < 16; ++n) carry[n] = out[16 + n];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Demodulation + output.  One 256-thread workgroup = kDemodLines field lines
// of one field of one frame.
// ---------------------------------------------------------------------------

constexpr int kDemodThreads = 256;            // = output columns per segment
constexpr int kSegCols = kDemodThreads + 2 * 16;  // staged columns incl. halo

struct LdsD {
    float chroma[kDemodLines + 6][kSegCols];
    float mn[8][kSegCols];
    float bp[kDemodLines], bq[kDemodLines], vsw[kDemodLines];
};

__device__ __forceinline__ float h273_round(float x) { return copysignf(floorf(fabsf(x) + 0.5f), x); }

// OutputWriter (output_writer.cpp:46-76, 124-190, 402-486, 670-777): the
// pixel format and the clamp box of the clamp mode, per launch.
struct OutSpec {
    int fmt;                              // chd_pixel_format_t
    int64_t plane;                        // samples per plane
    float eyLo, eyHi, ecLo, ecHi;         // float Y'CbCr
    float yLo, yHi, cLo, cHi;             // u16 Y'CbCr (H.273 narrow range)
};

__device__ __forceinline__ OutSpec out_spec(const Geom &g)
{
    OutSpec o;
    o.fmt = g.pixel_format;
    o.plane = (int64_t)g.out_w * g.out_h;
    o.eyLo = -INFINITY; o.eyHi = INFINITY; o.ecLo = -INFINITY; o.ecHi = INFINITY;
    o.yLo = 0.0f; o.yHi = 65535.0f; o.cLo = 0.0f; o.cHi = 65535.0f;
    if (g.clamp_mode == 1) {
        o.eyLo = 0.0f; o.eyHi = 1.0f; o.ecLo = -0.5f; o.ecHi = 0.5f;
        o.yLo = 16.0f * 256.0f; o.yHi = 235.0f * 256.0f; o.cLo = 16.0f * 256.0f; o.cHi = 240.0f * 256.0f;
    } else if (g.clamp_mode == 3) {
        o.eyLo = -15.0f / 219.0f; o.eyHi = 238.75f / 219.0f; o.ecLo = -127.0f / 224.0f; o.ecHi = 126.75f / 224.0f;
        o.yLo = 256.0f; o.yHi = 254.75f * 256.0f; o.cLo = 256.0f; o.cHi = 254.75f * 256.0f;
    }
    return o;
}

// One output pixel o (row-major index within a plane) from the
// composite-domain Y/U/V of the demodulator.
__device__ __forceinline__ void store_pixel(const Geom &g, const OutSpec &os, void *outp, int64_t o, float Y, float U,
                                            float V)
{
    const float ey = (Y - g.black) * g.ey_scale;
    if (os.fmt == 1 /*YUV444PS*/ || os.fmt == 5 /*GRAYS*/) {
        gfloat *out = (gfloat *)outp;
        out[o] = fminf(fmaxf(ey, os.eyLo), os.eyHi);
        if (os.fmt == 1) {
            out[os.plane + o] = fminf(fmaxf(U * g.ecb_scale, os.ecLo), os.ecHi);
            out[2 * os.plane + o] = fminf(fmaxf(V * g.ecr_scale, os.ecLo), os.ecHi);
        }
    } else if (os.fmt == 0 /*YUV444P16*/ || os.fmt == 4 /*GRAY16*/) {
        gu16 *out = (gu16 *)outp;
        out[o] = (uint16_t)fminf(fmaxf(h273_round(219.0f * 256.0f * ey + 16.0f * 256.0f), os.yLo), os.yHi);
        if (os.fmt == 0) {
            const float ecb = U * g.ecb_scale, ecr = V * g.ecr_scale;
            out[os.plane + o] = (uint16_t)fminf(fmaxf(h273_round(224.0f * 256.0f * ecb + 128.0f * 256.0f), os.cLo), os.cHi);
            out[2 * os.plane + o] = (uint16_t)fminf(fmaxf(h273_round(224.0f * 256.0f * ecr + 128.0f * 256.0f), os.cLo), os.cHi);
        }
    } else if (os.fmt == 3 /*RGBS*/) {
        // E'R/E'G/E'B planes (convertToFloatRGB)
        const float u = U * g.ey_scale, v = V * g.ey_scale;
        gfloat *out = (gfloat *)outp;
        out[o] = fminf(fmaxf(ey + g.rgb[0] * v, g.rgb_lo[0]), g.rgb_hi[0]);
        out[os.plane + o] = fminf(fmaxf(ey + g.rgb[1] * u + g.rgb[2] * v, g.rgb_lo[1]), g.rgb_hi[1]);
        out[2 * os.plane + o] = fminf(fmaxf(ey + g.rgb[3] * u, g.rgb_lo[2]), g.rgb_hi[2]);
    } else {
        // RGB48: interleaved full-range u16, luma clamped first, every
        // component clamped to [0, 65535] and truncated (convertLine)
        const float s = 65535.0f * g.ey_scale;
        const float ry = fminf(fmaxf((Y - g.black) * s, 0.0f), 65535.0f);
        const float ru = U * s, rv = V * s;
        gu16 *out = (gu16 *)outp + 3 * o;
        out[0] = (uint16_t)fminf(fmaxf(ry + g.rgb[0] * rv, 0.0f), 65535.0f);
        out[1] = (uint16_t)fminf(fmaxf(ry + g.rgb[1] * ru + g.rgb[2] * rv, 0.0f), 65535.0f);
        out[2] = (uint16_t)fminf(fmaxf(ry + g.rgb[3] * ru, 0.0f), 65535.0f);
    }
}

// Luma coring (PalColour::doYNR) needs the whole decoded E'Y line, so with
// it on the demodulators leave raw Y/U/V (composite units) in `raw`: per
// frame of a launch 3 planes of (last - first + 1) frame lines x aw samples.
__device__ __forceinline__ int64_t raw_frame_floats(const Geom &g) { return 3 * (int64_t)(g.last - g.first + 1) * g.aw; }

// Output pixel (frame_line, column c from active_video_start) of frame job
// `job_id`: converted, or (RAW) left uncored for luma coring.
template <bool RAW>
__device__ __forceinline__ void emit_pixel(const Geom &g, const OutSpec &os, const FrameJob &job, float *raw,
                                           int job_id, int frame_line, int c, float Y, float U, float V)
{
    if constexpr (RAW) {
        const int64_t plane = (int64_t)(g.last - g.first + 1) * g.aw;
        gfloat *r = (gfloat *)(raw + job_id * raw_frame_floats(g)) + (int64_t)(frame_line - g.first) * g.aw + c;
        r[0] = Y;
        r[plane] = U;
        r[2 * plane] = V;
    } else {
        const int64_t o = (int64_t)(frame_line - g.first + g.pad_top) * g.out_w + g.pad_left + c;
        store_pixel(g, os, job.out, o, Y, U, V);
    }
}

// Grid: jobs x 2 fields x line blocks x column segments.
__global__ void __launch_bounds__(kDemodThreads)
pal_demod(const Geom g, const FrameJob *__restrict__ jobs, const float *__restrict__ sine,
          const float *__restrict__ cosine, const float *__restrict__ cfilt /* [kMaxFilter+1][4] */,
          int n_lblocks, int n_segs, float *raw)
{
    __shared__ LdsD L;
    const int tid = threadIdx.x;
    int rem = blockIdx.x;
    const int seg = rem % n_segs; rem /= n_segs;
    const int lb = rem % n_lblocks; rem /= n_lblocks;
    const int fsel = rem & 1;
    const int job_id = rem >> 1;
    const FrameJob &job = jobs[job_id];
    const int off = job.offset[fsel];
    // two fields on the same lines: the second one's decode replaces the
    // first's (palcolour.cpp:295-296, 616)
    if (fsel == 0 && job.offset[1] == off) return;
    const uint16_t *comp = job.field[fsel];
    const int firstLine = (g.first + 1 - off) / 2;
    const int lastLine = ((g.last - off) / 2) + 1;
    const int n0 = firstLine + lb * kDemodLines;
    if (n0 >= lastLine) return;
    const int nl = min(kDemodLines, lastLine - n0);
    const int fs = g.filter_size;
    // PalColour::decodeLine fill window (palcolour.cpp:555-556)
    const int fillStart = max(0, g.avs - fs);
    const int fillEnd = min(g.ave + fs + 1, g.fw);
    const int endPos = min(g.ave, g.fw);
    // this segment's output columns and its staged window (halo 16 >= fs)
    const int i0 = g.avs + seg * kDemodThreads;
    const int s0 = i0 - 16;  // staged column 0

    // ---- stage chroma rows n0-3 .. n0+nl+2 (zero outside the active area) ---
    const int nrows = nl + 6;
    for (int idx = tid; idx < nrows * kSegCols; idx += kDemodThreads) {
        const int r = idx / kSegCols, cc = idx - r * kSegCols;
        const int line = n0 - 3 + r;
        const int col = s0 + cc;
        float v = 0.0f;
        const int nc = job.nchroma;
        if (line >= firstLine && line < lastLine) {
            if (nc == 0) {
                // CHD_DEC_PAL_2D: the filter input is the composite itself
                // (decodeLine<uint16_t, false>), real samples in the halo too
                if (col >= 0 && col < g.fw)
                    v = comp != nullptr ? (float)comp[(int64_t)line * g.pitch + col + g.lmargin] : g.black;
            } else if (col >= g.avs && col < g.ave) {
                const int64_t o = (int64_t)line * g.cpitch + (col - g.avs + kColOrigin);
                v = job.chroma[fsel][0][o];
                if (nc > 1) v += job.chroma[fsel][1][o];
                if (nc > 2) v += job.chroma[fsel][2][o];
                if (nc > 3) v += job.chroma[fsel][3][o];
            }
        }
        L.chroma[r][cc] = v;
    }
    const bool pal2d = job.nchroma == 0;

    // ---- burst detection per line (palcolour.cpp:344-409) -------------------
    {
        const int wave = tid >> 6, lane = tid & 63;
        const int blen = g.cbe - g.cbs;
        for (int li = wave; li < nl; li += kDemodThreads / 64) {
            const int number = n0 + li;
            float bp = 0, bq = 0, bpo = 0, bqo = 0;
            for (int i = g.cbs + lane; i < g.cbe; i += 64) {
                auto px = [&](int ln) -> float {
                    if (ln < 0 || ln >= g.fh) return 0.0f;  // blackLine (zeros)
                    if (comp == nullptr) return g.black;
                    return (float)((const gu16 *)comp)[(int64_t)ln * g.pitch + i + g.lmargin];
                };
                const float in0 = px(number), in1 = px(number - 1), in2 = px(number + 1);
                const float in3 = px(number - 2), in4 = px(number + 2);
                const float a = (in0 - (in3 + in4) * 0.5f) * 0.5f;
                const float b = (in2 - in1) * 0.5f;
                bp += a * sine[i];
                bq += a * cosine[i];
                bpo += b * sine[i];
                bqo += b * cosine[i];
            }
            bp = wave_total(bp) / blen;
            bq = wave_total(bq) / blen;
            bpo = wave_total(bpo) / blen;
            bqo = wave_total(bqo) / blen;
            if (lane == 0) {
                const float vs = (((bp - bpo) * (bp - bpo) + (bq - bqo) * (bq - bqo)) < (bp * bp + bq * bq) * 2) ? 1.0f : -1.0f;
                float lbp = (bp - bqo) * 0.5f, lbq = (bq + bpo) * 0.5f;
                const float norm = fmaxf(sqrtf(lbp * lbp + lbq * lbq), g.burst_floor);
                lbp /= norm;
                lbq /= norm;
                // gain / phase adjustment (palcolour.cpp:321-325)
                L.bp[li] = (lbp * g.chroma_gain_cos - lbq * g.chroma_gain_sin);
                L.bq[li] = (lbp * g.chroma_gain_sin + lbq * g.chroma_gain_cos);
                L.vsw[li] = vs;
            }
        }
    }
    __syncthreads();

    const OutSpec os = out_spec(g);

    const int i = i0 + tid;              // this thread's output column
    const bool col_live = i < endPos;
    for (int li = 0; li < nl; ++li) {
        const int number = n0 + li;
        const int r0 = li + 3;  // staged row of line `number`
        // m/n products (palcolour.cpp:557-567) over the staged columns
        for (int cc = tid; cc < kSegCols; cc += kDemodThreads) {
            const int col = s0 + cc;
            float S = 0.0f, Cn = 0.0f;
            if (col >= fillStart && col < fillEnd) { S = sine[col]; Cn = cosine[col]; }
            const float c0 = L.chroma[r0][cc];
            const float c1 = L.chroma[r0 - 1][cc], c2 = L.chroma[r0 + 1][cc];
            const float c3 = L.chroma[r0 - 2][cc], c4 = L.chroma[r0 + 2][cc];
            const float c5 = L.chroma[r0 - 3][cc], c6 = L.chroma[r0 + 3][cc];
            L.mn[0][cc] = c0 * S;
            L.mn[2][cc] = c1 * S - c2 * S;
            L.mn[1][cc] = -c3 * S - c4 * S;
            L.mn[3][cc] = -c5 * S + c6 * S;
            L.mn[4][cc] = c0 * Cn;
            L.mn[6][cc] = c1 * Cn - c2 * Cn;
            L.mn[5][cc] = -c3 * Cn - c4 * Cn;
            L.mn[7][cc] = -c5 * Cn + c6 * Cn;
        }
        __syncthreads();
        if (col_live) {
            const float bp = L.bp[li], bq = L.bq[li], vsw = L.vsw[li];
            float PU = 0, QU = 0, PV = 0, QV = 0, PY = 0, QY = 0;
            for (int b = 0; b <= fs; ++b) {
                // taps clamp into the populated halo (palcolour.cpp:587-588)
                const int l = max(fillStart, i - b) - s0;
                const int r = min(fillEnd - 1, i + b) - s0;
                const float c0 = cfilt[b * kFiltStride + 0], c1 = cfilt[b * kFiltStride + 1];
                const float c2 = cfilt[b * kFiltStride + 2], c3 = cfilt[b * kFiltStride + 3];
                const float m0 = L.mn[0][r] + L.mn[0][l], m1 = L.mn[1][r] + L.mn[1][l];
                const float m2 = L.mn[2][r] + L.mn[2][l], m3 = L.mn[3][r] + L.mn[3][l];
                const float q0 = L.mn[4][r] + L.mn[4][l], q1 = L.mn[5][r] + L.mn[5][l];
                const float q2 = L.mn[6][r] + L.mn[6][l], q3 = L.mn[7][r] + L.mn[7][l];
                const float common_p = m0 * c0 + m1 * c1;
                const float common_q = q0 * c0 + q1 * c1;
                PU += common_p + q2 * c2 + q3 * c3;
                QU += common_q - m2 * c2 - m3 * c3;
                PV += common_p - q2 * c2 - q3 * c3;
                QV += common_q + m2 * c2 + m3 * c3;
                if (pal2d) {
                    const float y0 = cfilt[b * kFiltStride + 4], y1 = cfilt[b * kFiltStride + 5];
                    PY += m0 * y0 + m1 * y1;
                    QY += q0 * y0 + q1 * y1;
                }
            }
            const float compv = (comp != nullptr) ? (float)comp[(int64_t)number * g.pitch + i + g.lmargin] : g.black;
            // luma: composite minus the pre-filtered chroma, or for PAL_2D minus
            // the chroma re-synthesised from the Y filter (palcolour.cpp:622-630)
            const float Y = pal2d ? compv - (PY * sine[i] + QY * cosine[i]) * 2.0f : compv - L.chroma[r0][i - s0];
            const float U = -(PU * bp + QU * bq) * 2.0f;
            const float V = vsw * -(QV * bp - PV * bq) * 2.0f;
            if (raw != nullptr) emit_pixel<true>(g, os, job, raw, job_id, number * 2 + off, i - g.avs, Y, U, V);
            else emit_pixel<false>(g, os, job, raw, job_id, number * 2 + off, i - g.avs, Y, U, V);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Demodulation at 4fSC (sampleRate = 4 fSC exactly, the ld-decode lattice):
// the reference carrier sin/cos(2 pi i / 4) is 0 or +-1, sine non-zero on
// columns of parity `spar` only and cosine on the others.  Every m[] term of
// decodeLine's FIR (palcolour.cpp:557-601) then vanishes on one column parity
// and every n[] term on the other; since taps i-b and i+b share a parity, the
// FIR regroups exactly into four symmetric convolutions of
//   t0 = c[n], t1 = -(c[n-2] + c[n+2]), t2 = c[n-1] - c[n+1], t3 = c[n+3] - c[n-3]
// (times the +-1 carrier), half the multiply-adds of the general form.
// One 256-thread workgroup = kQLines field lines x 256 output columns.
// ---------------------------------------------------------------------------

constexpr int kQLines = 16;
constexpr int kQHalo = 8;                             // >= filter_size
constexpr int kQCols = kDemodThreads + 2 * kQHalo;    // staged columns

struct LdsQ {
    float chroma[kQLines + 6][kQCols];
    float t[4][kQCols];
    float bp[kQLines], bq[kQLines], vsw[kQLines];
};

__global__ void __launch_bounds__(kDemodThreads)
pal_demod_quad(const Geom g, const FrameJob *__restrict__ jobs, const float *__restrict__ sine,
               const float *__restrict__ cosine, const float *__restrict__ cfilt, int n_lblocks, int n_segs,
               float *raw)
{
    __shared__ LdsQ L;
    const int tid = threadIdx.x;
    int rem = blockIdx.x;
    const int seg = rem % n_segs; rem /= n_segs;
    const int lb = rem % n_lblocks; rem /= n_lblocks;
    const int fsel = rem & 1;
    const int job_id = rem >> 1;
    const FrameJob &job = jobs[job_id];
    const int off = job.offset[fsel];
    // two fields on the same lines: the second one's decode replaces the
    // first's (palcolour.cpp:295-296, 616)
    if (fsel == 0 && job.offset[1] == off) return;
    const uint16_t *comp = job.field[fsel];
    const int firstLine = (g.first + 1 - off) / 2;
    const int lastLine = ((g.last - off) / 2) + 1;
    const int n0 = firstLine + lb * kQLines;
    if (n0 >= lastLine) return;
    const int nl = min(kQLines, lastLine - n0);
    const int fs = g.filter_size;
    const int i0 = g.avs + seg * kDemodThreads;
    const int s0 = i0 - kQHalo;

    // ---- stage chroma rows n0-3 .. n0+nl+2 (zero outside the active area) ---
    const int nrows = nl + 6;
    for (int idx = tid; idx < nrows * kQCols; idx += kDemodThreads) {
        const int r = idx / kQCols, cc = idx - r * kQCols;
        const int line = n0 - 3 + r;
        const int col = s0 + cc;
        float v = 0.0f;
        if (line >= firstLine && line < lastLine && col >= g.avs && col < g.ave) {
            const int64_t o = (int64_t)line * g.cpitch + (col - g.avs + kColOrigin);
            const int nc = job.nchroma;
            v = job.chroma[fsel][0][o];
            if (nc > 1) v += job.chroma[fsel][1][o];
            if (nc > 2) v += job.chroma[fsel][2][o];
            if (nc > 3) v += job.chroma[fsel][3][o];
        }
        L.chroma[r][cc] = v;
    }

    // ---- burst detection per line (palcolour.cpp:344-409) -------------------
    {
        const int wave = tid >> 6, lane = tid & 63;
        const int blen = g.cbe - g.cbs;
        for (int li = wave; li < nl; li += kDemodThreads / 64) {
            const int number = n0 + li;
            float bp = 0, bq = 0, bpo = 0, bqo = 0;
            for (int i = g.cbs + lane; i < g.cbe; i += 64) {
                auto px = [&](int ln) -> float {
                    if (ln < 0 || ln >= g.fh) return 0.0f;  // blackLine (zeros)
                    if (comp == nullptr) return g.black;
                    return (float)((const gu16 *)comp)[(int64_t)ln * g.pitch + i + g.lmargin];
                };
                const float in0 = px(number), in1 = px(number - 1), in2 = px(number + 1);
                const float in3 = px(number - 2), in4 = px(number + 2);
                const float a = (in0 - (in3 + in4) * 0.5f) * 0.5f;
                const float b = (in2 - in1) * 0.5f;
                bp += a * sine[i];
                bq += a * cosine[i];
                bpo += b * sine[i];
                bqo += b * cosine[i];
            }
            bp = wave_total(bp) / blen;
            bq = wave_total(bq) / blen;
            bpo = wave_total(bpo) / blen;
            bqo = wave_total(bqo) / blen;
            if (lane == 0) {
                const float vs = (((bp - bpo) * (bp - bpo) + (bq - bqo) * (bq - bqo)) < (bp * bp + bq * bq) * 2) ? 1.0f : -1.0f;
                float lbp = (bp - bqo) * 0.5f, lbq = (bq + bpo) * 0.5f;
                const float norm = fmaxf(sqrtf(lbp * lbp + lbq * lbq), g.burst_floor);
                lbp /= norm;
                lbq /= norm;
                L.bp[li] = (lbp * g.chroma_gain_cos - lbq * g.chroma_gain_sin);
                L.bq[li] = (lbp * g.chroma_gain_sin + lbq * g.chroma_gain_cos);
                L.vsw[li] = vs;
            }
        }
    }
    __syncthreads();

    const OutSpec os = out_spec(g);
    float cf[kQHalo + 1][4];
#pragma unroll
    for (int b = 0; b <= kQHalo; ++b)
#pragma unroll
        for (int j = 0; j < 4; ++j) cf[b][j] = (b <= fs) ? cfilt[b * kFiltStride + j] : 0.0f;

    const int i = i0 + tid;
    const bool col_live = i < g.ave;
    const int ci = i - s0;
    for (int li = 0; li < nl; ++li) {
        const int number = n0 + li;
        const int r0 = li + 3;
        for (int cc = tid; cc < kQCols; cc += kDemodThreads) {
            const int col = s0 + cc;
            // the carrier: exactly one of sine/cosine is non-zero (+-1)
            const float w = (col >= 0 && col < g.fw) ? sine[col] + cosine[col] : 0.0f;
            const float c0 = L.chroma[r0][cc];
            L.t[0][cc] = c0 * w;
            L.t[1][cc] = -(L.chroma[r0 - 2][cc] + L.chroma[r0 + 2][cc]) * w;
            L.t[2][cc] = (L.chroma[r0 - 1][cc] - L.chroma[r0 + 1][cc]) * w;
            L.t[3][cc] = (L.chroma[r0 + 3][cc] - L.chroma[r0 - 3][cc]) * w;
        }
        __syncthreads();
        if (col_live) {
            float P01 = 0.0f, Q01 = 0.0f, P23 = 0.0f, Q23 = 0.0f;
#pragma unroll
            for (int b = 0; b <= kQHalo; ++b) {
                const float s0_ = L.t[0][ci + b] + L.t[0][ci - b];
                const float s1_ = L.t[1][ci + b] + L.t[1][ci - b];
                const float s2_ = L.t[2][ci + b] + L.t[2][ci - b];
                const float s3_ = L.t[3][ci + b] + L.t[3][ci - b];
                const float a01 = fmaf(s1_, cf[b][1], s0_ * cf[b][0]);
                const float a23 = fmaf(s3_, cf[b][3], s2_ * cf[b][2]);
                if (((i + b) & 1) == g.spar) {  // sine column: m terms
                    P01 += a01;
                    Q23 += a23;
                } else {                          // cosine column: n terms
                    Q01 += a01;
                    P23 += a23;
                }
            }
            const float PU = P01 + P23, PV = P01 - P23, QU = Q01 - Q23, QV = Q01 + Q23;
            const float bp = L.bp[li], bq = L.bq[li], vsw = L.vsw[li];
            const float compv = (comp != nullptr) ? (float)comp[(int64_t)number * g.pitch + i + g.lmargin] : g.black;
            const float Y = compv - L.chroma[r0][ci];
            const float U = -(PU * bp + QU * bq) * 2.0f;
            const float V = vsw * -(QV * bp - PV * bq) * 2.0f;
            if (raw != nullptr) emit_pixel<true>(g, os, job, raw, job_id, number * 2 + off, i - g.avs, Y, U, V);
            else emit_pixel<false>(g, os, job, raw, job_id, number * 2 + off, i - g.avs, Y, U, V);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 4fSC demodulation, column-blocked.  One 256-thread workgroup = kQ4Lines
// field lines x the whole active line; thread t owns the 4 columns
// 4t - 8 .. 4t - 5 (from active_video_start), so threads 0-1 hold the left
// filter halo.  The thread keeps the 7 input rows its columns need (line
// n-3 .. n+3) in registers and slides them down the block's lines, so each
// partial plane is read once plus a 6-row halo per block, as float4.  The
// FIR inputs t0..t3 of a line (the same definitions and summation order as
// pal_demod_quad) pass through LDS once, double-buffered: one barrier per
// line.  NC is the launch's FrameJob.nchroma: 4 (3D), 2 (2D) or 1 partial
// planes summed per row, or 0 (CHD_DEC_PAL_2D), which filters the composite
// itself and takes luma from the re-synthesised chroma.  A row's loads are
// issued lines ahead and summed only after the current line's filter, so
// HBM latency overlaps lines of work.
// ---------------------------------------------------------------------------

constexpr int kQ4Lines = 48;
constexpr int kQ4ShortLines = 8;                       // lines per workgroup of small launches
constexpr int kQ4Cols = 4 * kDemodThreads;             // columns held by the threads
constexpr int kQ4MaxWidth = kQ4Cols - 2 * kQHalo;       // widest active line: halos on both sides
constexpr int kQ4Row4 = kDemodThreads + 2;              // float4 per t row (+ 8 zero columns)

template <int LINES>
struct LdsQ4 {
    float4 t[2][4][kQ4Row4];
    float bp[LINES], bq[LINES], vsw[LINES];
};

__device__ __forceinline__ float4 f4add(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ float f4get(const float4 &v, int e) { return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w; }

// lpb: lines per workgroup (at most LINES), chosen by the launcher so that
// the workgroups fill whole rounds of the resident slots.
// The 4-plane (Transform-3D) form is held to 3 waves per SIMD (168 VGPRs,
// two of them spilled): left free it takes 187 and runs at 2 waves, 3% slower.
template <int NC, bool RAW, int LINES>
__global__ void __launch_bounds__(kDemodThreads, NC == 4 ? 3 : 1)
pal_demod_quad4(const Geom g, const FrameJob *__restrict__ jobs, const float *__restrict__ sine,
                const float *__restrict__ cosine, const float *__restrict__ cfilt, int n_lblocks, int lpb, float *raw)
{
    __shared__ LdsQ4<LINES> L;
    const int tid = threadIdx.x;
    const int lb = blockIdx.x % n_lblocks;
    const int fsel = (blockIdx.x / n_lblocks) & 1;
    const int job_id = blockIdx.x / (2 * n_lblocks);
    const FrameJob &job = jobs[job_id];
    const int off = job.offset[fsel];
    // two fields on the same lines: the second one's decode replaces the
    // first's (palcolour.cpp:295-296, 616)
    if (fsel == 0 && job.offset[1] == off) return;
    const uint16_t *comp = job.field[fsel];
    const int firstLine = (g.first + 1 - off) / 2;
    const int lastLine = ((g.last - off) / 2) + 1;
    const int n0 = firstLine + lb * lpb;
    if (n0 >= lastLine) return;
    const int nl = min(lpb, lastLine - n0);
    const int fs = g.filter_size;
    const int c0 = 4 * tid - kQHalo;         // this thread's first column, from active_video_start
    constexpr bool pal2d = NC == 0;
    // columns whose input this thread loads: the active line for separated
    // chroma (zero elsewhere), the line plus the filter halo for PAL_2D
    const bool loads = pal2d ? (c0 + 4 > -kQHalo && c0 < g.aw + kQHalo) : (c0 >= 0 && c0 < g.aw);
    const bool outputs = c0 >= 0 && c0 < g.aw;
    const float *const *planes = job.chroma[fsel];

    // zero columns beyond the threads' range (both buffers, all 4 rows)
    if (tid < 16) {
        const int buf = tid >> 3, k = (tid >> 1) & 3, h = tid & 1;
        L.t[buf][k][kDemodThreads + h] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }

    // ---- burst detection per line (palcolour.cpp:344-409) -------------------
    // A wave takes lines li = wave, wave + 4, ... (at most LINES / 4), in
    // batches of kQ lines whose loads are all issued before any is summed:
    // lines n, n-1, n+1, n-2 and n+2 of each, every loaded value used (times
    // 0 where it does not count) so that none moves under a branch of its
    // own.  A line outside the field is the black line of zeros, an absent
    // field black (its loads read the filter table).
    {
        const int wave = tid >> 6, lane = tid & 63;
        const int blen = g.cbe - g.cbs;
        constexpr int kPer = LINES / (kDemodThreads / 64);
        constexpr int kQ = kPer < 2 ? kPer : 2;
        constexpr int kDl[5] = {0, -1, 1, -2, 2};
        const uint16_t *const src = comp != nullptr ? comp : reinterpret_cast<const uint16_t *>(cfilt);
        for (int l0 = wave; l0 < nl; l0 += 4 * kQ) {
        float bp[kQ], bq[kQ], bpo[kQ], bqo[kQ];
#pragma unroll
        for (int q = 0; q < kQ; ++q) bp[q] = bq[q] = bpo[q] = bqo[q] = 0.0f;
        for (int i = g.cbs + lane; i < g.cbe; i += 64) {
            const float sn = sine[i], cs = cosine[i];
            float raw[kQ][5];
#pragma unroll
            for (int q = 0; q < kQ; ++q)
#pragma unroll
                for (int d = 0; d < 5; ++d) {
                    const int ln = min(max(n0 + l0 + 4 * q + kDl[d], 0), g.fh - 1);
                    const int64_t o = comp != nullptr ? (int64_t)ln * g.pitch + i + g.lmargin : 0;
                    raw[q][d] = (float)((const gu16 *)src)[o];
                }
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                float in[5];
#pragma unroll
                for (int d = 0; d < 5; ++d) {
                    const int ln = n0 + l0 + 4 * q + kDl[d];
                    const bool inside = ln >= 0 && ln < g.fh;
                    in[d] = raw[q][d] * (inside && comp != nullptr ? 1.0f : 0.0f) +
                            (inside && comp == nullptr ? g.black : 0.0f);
                }
                const float a = (in[0] - (in[3] + in[4]) * 0.5f) * 0.5f;
                const float b = (in[2] - in[1]) * 0.5f;
                bp[q] += a * sn;
                bq[q] += a * cs;
                bpo[q] += b * sn;
                bqo[q] += b * cs;
            }
        }
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int li = l0 + 4 * q;
            if (li >= nl) break;
            const float tp = wave_total(bp[q]) / blen, tq = wave_total(bq[q]) / blen;
            const float tpo = wave_total(bpo[q]) / blen, tqo = wave_total(bqo[q]) / blen;
            if (lane == 0) {
                const float vs =
                    (((tp - tpo) * (tp - tpo) + (tq - tqo) * (tq - tqo)) < (tp * tp + tq * tq) * 2) ? 1.0f : -1.0f;
                float lbp = (tp - tqo) * 0.5f, lbq = (tq + tpo) * 0.5f;
                const float norm = fmaxf(sqrtf(lbp * lbp + lbq * lbq), g.burst_floor);
                lbp /= norm;
                lbq /= norm;
                L.bp[li] = (lbp * g.chroma_gain_cos - lbq * g.chroma_gain_sin);
                L.bq[li] = (lbp * g.chroma_gain_sin + lbq * g.chroma_gain_cos);
                L.vsw[li] = vs;
            }
        }
        }
    }

    // Filter input of this thread's columns on field line `line`, in two
    // steps: issue() starts the loads, sum() forms the value.  It is the
    // partial planes summed in order (zero outside the active area), or the
    // composite (PAL_2D: black when the field is absent); zero on lines
    // outside the field's active lines.
    // The loads are unconditional, so that all of them sit in one basic
    // block and the compiler's wait counts stay exact: a lane with nothing
    // to load reads the filter table instead and drops the value in sum().
    struct Raw {
        float4 p[NC > 0 ? NC : 1];
        uint2 c;
        bool ok;
    };
    const uint16_t *const dummy16 = reinterpret_cast<const uint16_t *>(cfilt);
    // this thread's column in each input, at line 0 (the dummy where the
    // thread loads nothing)
    const float *pbase[NC > 0 ? NC : 1];
#pragma unroll
    for (int k = 0; k < NC; ++k) pbase[k] = loads ? planes[k] + kColOrigin + c0 : cfilt;
    const uint16_t *const cbase = (loads && comp != nullptr) ? comp + g.lmargin + g.avs + c0 : dummy16;
    auto issue = [&](int line) -> Raw {
        Raw w;
        w.ok = loads && line >= firstLine && line < lastLine;
        const int64_t row = w.ok ? line : 0;
        if constexpr (pal2d) {
            w.c = gload_u2(cbase + (cbase != dummy16 ? row * g.pitch : 0));
        } else {
            const int64_t o = (pbase[0] != cfilt) ? row * g.cpitch : 0;
#pragma unroll
            for (int k = 0; k < NC; ++k) w.p[k] = gload_f4(pbase[k] + o);
        }
        return w;
    };
    auto sum = [&](const Raw &w) -> float4 {
        const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if constexpr (pal2d) {
            if (!w.ok) return zero;
            if (comp == nullptr) return make_float4(g.black, g.black, g.black, g.black);
            return make_float4((float)(w.c.x & 0xFFFFu), (float)(w.c.x >> 16), (float)(w.c.y & 0xFFFFu),
                               (float)(w.c.y >> 16));
        } else {
            float4 v = w.p[0];
#pragma unroll
            for (int k = 1; k < NC; ++k) v = f4add(v, w.p[k]);
            if (!w.ok) v = zero;
            if (c0 + 4 > g.aw) {
                if (c0 + 1 >= g.aw) v.y = 0.0f;
                if (c0 + 2 >= g.aw) v.z = 0.0f;
                v.w = 0.0f;
            }
            return v;
        }
    };

    // carrier (exactly one of sine/cosine is non-zero, +-1) of the 4 columns
    // (every load issued before any is used and every value used: a column
    // outside the field reads a clamped column and is weighted 0)
    float w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int col = g.avs + c0 + e;
        const int cc = min(max(col, 0), g.fw - 1);
        w[e] = (sine[cc] + cosine[cc]) * ((loads && col >= 0 && col < g.fw) ? 1.0f : 0.0f);
    }
    constexpr int NCF = pal2d ? 6 : 4;   // chroma taps, plus the luma taps for PAL_2D
    float cf[kQHalo + 1][NCF];
#pragma unroll
    for (int b = 0; b <= kQHalo; ++b)
#pragma unroll
        for (int j = 0; j < NCF; ++j) cf[b][j] = (b <= fs) ? cfilt[b * kFiltStride + j] : 0.0f;
    // columns c0 + e + b of the sine parity: ((e + b) & 1) == sel for every thread
    const int sel = g.spar ^ ((g.avs + c0) & 1);

    const OutSpec os = out_spec(g);

    // rows n0-3 .. n0+3 of this thread's columns, each row's loads issued
    // three rows ahead of its sum (then line n0+4's, for the loop)
    float4 r[7];
    Raw pre[3], qa;
    pre[0] = issue(n0 - 3);
    pre[1] = issue(n0 - 2);
    pre[2] = issue(n0 - 1);
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        r[k] = sum(pre[k % 3]);
        if (k + 3 < 7) pre[k % 3] = issue(n0 + k);
        else if (k + 3 == 7) qa = issue(n0 + 4);
    }
    // Line li: rows r[0..6] are lines li-3 .. li+3; `use` holds the loads
    // of line li+4, `refill` receives those of line li+5.  The loop is
    // unrolled by two with the two buffers swapping roles, so that no load
    // result is copied (a copy would wait for it).
    auto line = [&](int li, const Raw &use, Raw &refill) {
        const int number = n0 + li;
        const int buf = li & 1;
        // this line's composite, then the row two lines ahead, in flight
        // through the filter (loads complete in order: the composite first)
        uint2 cw = make_uint2(0u, 0u);
        if constexpr (!pal2d) cw = gload_u2(cbase + (cbase != dummy16 ? (int64_t)number * g.pitch : 0));
        refill = issue(number + 5);
        {
            float4 t[4];
            float *t0 = &t[0].x, *t1 = &t[1].x, *t2 = &t[2].x, *t3 = &t[3].x;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                t0[e] = f4get(r[3], e) * w[e];
                t1[e] = -(f4get(r[1], e) + f4get(r[5], e)) * w[e];
                t2[e] = (f4get(r[2], e) - f4get(r[4], e)) * w[e];
                t3[e] = (f4get(r[6], e) - f4get(r[0], e)) * w[e];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) L.t[buf][k][tid] = t[k];
        }
        __syncthreads();
        if (outputs) {
            // FIR over t_k at columns c0-8 .. c0+11 (float4 tid-2 .. tid+2), in
            // two passes (t0,t1 then t2,t3) to halve the live registers; per
            // output the additions keep pal_demod_quad's order.
            float A01[2][4], A23[2][4], AY[2][4];   // [tap parity (e + b) & 1][column e]
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                float T[2][20];
#pragma unroll
                for (int k = 0; k < 2; ++k)
#pragma unroll
                    for (int m = 0; m < 5; ++m) {
                        const float4 q = L.t[buf][2 * half + k][tid - 2 + m];
                        T[k][4 * m] = q.x; T[k][4 * m + 1] = q.y; T[k][4 * m + 2] = q.z; T[k][4 * m + 3] = q.w;
                    }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float ev = 0.0f, od = 0.0f, yev = 0.0f, yod = 0.0f;
#pragma unroll
                    for (int b = 0; b <= kQHalo; ++b) {
                        const int ci = 8 + e;
                        const float sa = T[0][ci + b] + T[0][ci - b];
                        const float sb = T[1][ci + b] + T[1][ci - b];
                        const float a = fmaf(sb, cf[b][2 * half + 1], sa * cf[b][2 * half]);
                        if (((e + b) & 1) == 0) ev += a;
                        else od += a;
                        if constexpr (pal2d) {
                            if (half == 0) {
                                // Y filter: lines 0 and +-2 (t0, t1)
                                const float ay = fmaf(sb, cf[b][5], sa * cf[b][4]);
                                if (((e + b) & 1) == 0) yev += ay;
                                else yod += ay;
                            }
                        }
                    }
                    (half == 0 ? A01 : A23)[0][e] = ev;
                    (half == 0 ? A01 : A23)[1][e] = od;
                    if (half == 0) {
                        AY[0][e] = yev;
                        AY[1][e] = yod;
                    }
                }
            }
            const float bp = L.bp[li], bq = L.bq[li], vsw = L.vsw[li];
            // the composite of this line: PAL_2D's filter input row itself
            float compv[4];
            if (pal2d) {
#pragma unroll
                for (int e = 0; e < 4; ++e) compv[e] = f4get(r[3], e);
            } else {
                // an absent field reads as black
                if (comp == nullptr) cw = make_uint2(0u, 0u);
                const float cb = comp != nullptr ? 0.0f : g.black;
                compv[0] = cb + (float)(cw.x & 0xFFFFu); compv[1] = cb + (float)(cw.x >> 16);
                compv[2] = cb + (float)(cw.y & 0xFFFFu); compv[3] = cb + (float)(cw.y >> 16);
            }
            const int frame_line = number * 2 + off;
            const int64_t obase = (int64_t)(frame_line - g.first + g.pad_top) * g.out_w + g.pad_left + c0;
            // taps with e + b of parity sel sit on sine columns (m terms: P01, Q23)
            const int ps = sel, pc = sel ^ 1;
            // yuv444ps pixel pairs on 8-byte boundaries (output buffer, this
            // line's offset and the plane size all even in floats) take one
            // 8-byte store per plane.  (Not in the 1-plane and composite
            // forms: there the second path costs them the fourth wave per
            // SIMD.)
            const bool pairs = !RAW && NC >= 2 && os.fmt == 1 /*YUV444PS*/ && c0 + 4 <= g.aw &&
                               ((obase | os.plane | (int64_t)(reinterpret_cast<uintptr_t>(job.out) >> 2)) & 1) == 0 &&
                               (reinterpret_cast<uintptr_t>(job.out) & 3) == 0;
#pragma unroll
            for (int e2 = 0; e2 < 4; e2 += 2) {
                float Yv[2], Uv[2], Vv[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int e = e2 + h;
                    const float P01 = A01[ps][e], Q01 = A01[pc][e], Q23 = A23[ps][e], P23 = A23[pc][e];
                    const float PU = P01 + P23, PV = P01 - P23, QU = Q01 - Q23, QV = Q01 + Q23;
                    // PAL_2D luma: composite minus (PY sine + QY cosine) * 2, and
                    // column e is itself a sine column when (e & 1) == sel
                    float Y;
                    if constexpr (pal2d) Y = compv[e] - ((e & 1) == sel ? AY[ps][e] : AY[pc][e]) * w[e] * 2.0f;
                    else Y = compv[e] - f4get(r[3], e);
                    Yv[h] = Y;
                    Uv[h] = -(PU * bp + QU * bq) * 2.0f;
                    Vv[h] = vsw * -(QV * bp - PV * bq) * 2.0f;
                }
                if (pairs) {
                    // the values store_pixel forms
                    float ey[2], cb[2], cr[2];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        ey[h] = fminf(fmaxf((Yv[h] - g.black) * g.ey_scale, os.eyLo), os.eyHi);
                        cb[h] = fminf(fmaxf(Uv[h] * g.ecb_scale, os.ecLo), os.ecHi);
                        cr[h] = fminf(fmaxf(Vv[h] * g.ecr_scale, os.ecLo), os.ecHi);
                    }
                    float *out = static_cast<float *>(job.out) + obase + e2;
                    gstore_f2(out, ey[0], ey[1]);
                    gstore_f2(out + os.plane, cb[0], cb[1]);
                    gstore_f2(out + 2 * os.plane, cr[0], cr[1]);
                } else {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int e = e2 + h;
                        if (c0 + e < g.aw) {
                            if constexpr (RAW) emit_pixel<true>(g, os, job, raw, job_id, frame_line, c0 + e, Yv[h], Uv[h], Vv[h]);
                            else store_pixel(g, os, job.out, obase + e, Yv[h], Uv[h], Vv[h]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) r[k] = r[k + 1];
        r[6] = sum(use);
    };
    Raw qb;
    for (int li = 0; li < nl; li += 2) {
        line(li, qa, qb);
        if (li + 1 < nl) line(li + 1, qb, qa);
    }
}

// ---------------------------------------------------------------------------
// Packing and border fill.
// ---------------------------------------------------------------------------

// Copy .tbc-layout fields (fh x fw uint16, contiguous) into the packed store.
// shift > 0 drops the first `shift` samples of every field and appends as many
// black ones (SourceField::shiftSecondField, source_field.cpp:51-66).
__global__ void pack_fields(const Geom g, const uint16_t *__restrict__ src, int64_t src_field_elems,
                            uint16_t *__restrict__ dst, int n_fields, int shift)
{
    const int64_t total = (int64_t)n_fields * g.fh * g.pitch;
    const int64_t samples = (int64_t)g.fh * g.fw;
    const uint16_t blk = (uint16_t)g.black;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = i / g.field_elems;
        const int64_t r = i - f * g.field_elems;
        const int line = (int)(r / g.pitch);
        const int col = (int)(r - (int64_t)line * g.pitch) - g.lmargin;
        const int64_t at = (int64_t)line * g.fw + col + shift;
        dst[i] = (col >= 0 && col < g.fw && at < samples) ? src[f * src_field_elems + at] : blk;
    }
}

// The same for a list of fields: staged field blockIdx.y goes to its own slot.
__global__ void pack_field_list(const Geom g, const uint16_t *__restrict__ src, int64_t src_field_elems,
                                uint16_t *__restrict__ dst, const PackList list)
{
    const int k = blockIdx.y;
    const uint16_t *f = src + (int64_t)k * src_field_elems;
    uint16_t *out = dst + (int64_t)list.slot[k] * g.field_elems;
    const int shift = list.shift[k];
    const int64_t samples = (int64_t)g.fh * g.fw;
    const uint16_t blk = (uint16_t)g.black;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < g.field_elems;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int line = (int)(r / g.pitch);
        const int col = (int)(r - (int64_t)line * g.pitch) - g.lmargin;
        const int64_t at = (int64_t)line * g.fw + col + shift;
        out[r] = (col >= 0 && col < g.fw && at < samples) ? f[at] : blk;
    }
}

// Bulk copy in 16-byte words; the last bytes of a length that is not a
// multiple of 16 go one by one.
__global__ void copy_in(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16,
                        const uint8_t *__restrict__ src_tail, uint8_t *__restrict__ dst_tail, int tail)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
    if (blockIdx.x == 0 && (int)threadIdx.x < tail) dst_tail[threadIdx.x] = src_tail[threadIdx.x];
}

// Chroma fields (field_height x field_width floats each) into the partial
// plane layout the demodulator reads (columns of the active samples only).
__global__ void chroma_to_planes(const Geom g, const float *__restrict__ src, float *__restrict__ dst,
                                 int64_t plane_stride, int n_fields)
{
    const int64_t per = (int64_t)g.fh * g.aw;
    const int64_t total = (int64_t)n_fields * per;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = i / per;
        const int64_t r = i - f * per;
        const int line = (int)(r / g.aw);
        const int c = (int)(r - (int64_t)line * g.aw);
        dst[f * plane_stride + (int64_t)line * g.cpitch + c + kColOrigin] =
            src[(f * g.fh + line) * (int64_t)g.fw + g.avs + c];
    }
}

// Fill whole output frames with the black border value (only needed when
// padding adds a border around the active picture).
__global__ void fill_border(const Geom g, void *out, int64_t n_frames)
{
    // black: 0.0 for float formats, 0 for RGB48, H.273 digital zero for Y'CbCr
    const bool is_float = (g.pixel_format == 1 || g.pixel_format == 3 || g.pixel_format == 5);
    const bool rgb48 = g.pixel_format == 2;
    const int planes = (g.pixel_format == 4 || g.pixel_format == 5) ? 1 : 3;
    const int64_t plane = (int64_t)g.out_w * g.out_h;
    const int64_t total = n_frames * planes * plane;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int p = (int)((i / plane) % planes);
        if (is_float) {
            reinterpret_cast<float *>(out)[i] = 0.0f;
        } else if (rgb48) {
            reinterpret_cast<uint16_t *>(out)[i] = 0;
        } else {
            reinterpret_cast<uint16_t *>(out)[i] = (p == 0) ? (uint16_t)(16 * 256) : (uint16_t)(128 * 256);
        }
    }
}

// Luma coring, PalColour::doYNR (palcolour.cpp:411-448): each decoded frame
// line's E'Y minus its high-pass clipped to +-g.ynr.  The high-pass is the
// 25-tap symmetric FIR f_nrpal (deemp.h:620-633) centred on the sample, over
// the active samples with zeros outside them.  kNrTaps[j] is tap 12 +- j.
constexpr float kNrTaps[13] = {
    7.479077367478024e-01f, -2.223459379380252e-01f, -1.492080690537196e-01f, -6.398130386469833e-02f,
    1.098965697652802e-03f, 3.029931283148555e-02f, 2.861211356202848e-02f, 1.309163063177155e-02f,
    -4.423694440267179e-04f, -5.577680979937061e-03f, -4.499636864042073e-03f, -1.857019211291029e-03f,
    1.141291975113614e-04f};

// Grid: frame jobs x (last - first + 1) frame lines; raw as the demodulators
// left it.  Lines of a parity no field was decoded into are skipped (the
// two-fields-on-one-parity fill writes them).
__global__ void __launch_bounds__(256) ynr_output(const Geom g, const FrameJob *__restrict__ jobs,
                                                  const float *__restrict__ raw)
{
    const int nlines = g.last - g.first + 1;
    const int r = blockIdx.x % nlines, job_id = blockIdx.x / nlines;
    const FrameJob &job = jobs[job_id];
    const int par = (g.first + r) & 1;
    if (par != job.offset[0] && par != job.offset[1]) return;
    const OutSpec os = out_spec(g);
    const int64_t plane = (int64_t)nlines * g.aw;
    const float *yl = raw + job_id * raw_frame_floats(g) + (int64_t)r * g.aw;
    for (int c = threadIdx.x; c < g.aw; c += blockDim.x) {
        float a = kNrTaps[0] * yl[c];
#pragma unroll
        for (int j = 1; j <= 12; ++j) {
            const float lv = c - j >= 0 ? yl[c - j] : 0.0f;
            const float rv = c + j < g.aw ? yl[c + j] : 0.0f;
            a = fmaf(kNrTaps[j], lv + rv, a);
        }
        a = fminf(fmaxf(a, -g.ynr), g.ynr);
        store_pixel(g, os, job.out, (int64_t)(r + g.pad_top) * g.out_w + g.pad_left + c, yl[c] - a, yl[plane + c],
                    yl[2 * plane + c]);
    }
}

// A frame whose two fields share one line offset leaves the active lines of
// the other parity undecoded: they hold the zero Y/U/V the component frame
// starts with (component_frame.cpp:47-52), converted like any pixel.
__global__ void fill_other_parity(const Geom g, void *out, int parity)
{
    const OutSpec os = out_spec(g);
    const int rows = (g.last - g.first + 2) / 2;  // >= the lines of one parity
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)rows * g.aw;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / g.aw), c = (int)(i - (int64_t)r * g.aw);
        const int line = g.first + ((g.first ^ parity) & 1) + 2 * r;
        if (line > g.last) continue;
        store_pixel(g, os, out, (int64_t)(line - g.first + g.pad_top) * g.out_w + g.pad_left + c, 0.0f, 0.0f, 0.0f);
    }
}

}  // namespace tpal

// Launch wrappers (C++ linkage, used by tpal_device.cpp).
namespace tpal {

hipError_t upload_tables(const Tables &t) { return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &t, sizeof(Tables)); }

// Tiles per work item when the caller leaves it open: a workgroup per item
// (12 resident per CU at the tile kernel's footprint), each item costing its
// owned tiles plus the leading one, so a launch takes about
// ceil(items / resident) * (strip + 1) tile times.  Longer strips waste less
// on leading tiles; shorter ones fill the chip when there are few slabs.
// Strips are multiples of 4 tiles, so every strip's partial columns start on
// a 128-byte line.
int choose_strip3(const Geom &g, int n_jobs)
{
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            return n;
        return 256;
    }();
    const int64_t resident = 12 * (int64_t)cus;
    int best = kStrip3;
    int64_t best_cost = INT64_MAX;
    for (int strip : {4, 8, 12, 16, 24}) {
        const int64_t items = (int64_t)n_jobs * g.n_ty3 * ((g.n_tx3 + strip - 1) / strip);
        const int64_t cost = (items + resident - 1) / resident * (strip + 1);
        if (cost < best_cost) {
            best_cost = cost;
            best = strip;
        }
    }
    return best;
}

void launch_tpal3d(const Geom &g, const SlabJob *jobs, int n_jobs, const float *thr32, const double *thr64,
                   hipStream_t s, int mode, void *spec, const float *mask, int strip, int max_groups, unsigned *wq,
                   bool half)
{
    float2 *sp = reinterpret_cast<float2 *>(spec);
    if (strip <= 0) strip = choose_strip3(g, n_jobs);
    strip = (strip + 3) & ~3;   // strips start on a 4-tile group (whole-line stores)
    const int items = n_jobs * g.n_ty3 * ((g.n_tx3 + strip - 1) / strip);
    if (items <= 0) return;
    // persistent grids are whole multiples of 8 (one share per XCD)
    const bool persist = max_groups > 0 && max_groups < items && wq != nullptr;
    const dim3 grid(persist ? std::max(8, max_groups & ~7) : items), blk(64);
#define CHD_LAUNCH3D(M, P, H)                                                                                \
    hipLaunchKernelGGL((tpal3d_tiles<M, P, H>), grid, blk, 0, s, g, jobs, n_jobs, strip, thr32, thr64, sp, mask, wq)
    if (mode == kTileFeatures || mode == kTileApplyFwd) {
        // the paired learned path: half form only (the host checked the slabs)
        if (!half) return;
        if (mode == kTileFeatures) {
            if (persist) CHD_LAUNCH3D(kTileFeatures, true, true); else CHD_LAUNCH3D(kTileFeatures, false, true);
        } else {
            if (persist) CHD_LAUNCH3D(kTileApplyFwd, true, true); else CHD_LAUNCH3D(kTileApplyFwd, false, true);
        }
    } else if (mode == kTileSpectra) {
        if (persist) CHD_LAUNCH3D(kTileSpectra, true, false); else CHD_LAUNCH3D(kTileSpectra, false, false);
    } else if (mode == kTileApply) {
        if (persist) CHD_LAUNCH3D(kTileApply, true, false); else CHD_LAUNCH3D(kTileApply, false, false);
    } else if (half && g.thr_uniform >= 0.0f) {
        if (persist) CHD_LAUNCH3D(kTilePlain, true, true); else CHD_LAUNCH3D(kTilePlain, false, true);
    } else {
        if (persist) CHD_LAUNCH3D(kTilePlain, true, false); else CHD_LAUNCH3D(kTilePlain, false, false);
    }
#undef CHD_LAUNCH3D
}

void launch_tpal2d(const Geom &g, const FieldJob2D *jobs, int n_jobs, int n_ty_max, const float *thr32,
                   const double *thr64, hipStream_t s)
{
    const int groups_per_row = (g.n_strips2 + 3) / 4;
    const int blocks = n_jobs * n_ty_max * groups_per_row;
    if (blocks > 0)
        hipLaunchKernelGGL(tpal2d_tiles, dim3(blocks), dim3(64), 0, s, g, jobs, n_ty_max, thr32, thr64);
}

// Compute units of the current device, and how many workgroups of a kernel
// one of them holds (both cached per device / kernel).
static int cu_count()
{
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cached[dev].load(std::memory_order_relaxed);
    if (n == 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cached[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

static int resident_per_cu(const void *kern)
{
    static std::mutex mu;
    static std::vector<std::pair<const void *, int>> cache;
    std::lock_guard<std::mutex> lock(mu);
    for (const auto &e : cache)
        if (e.first == kern) return e.second;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, kDemodThreads, 0) != hipSuccess || n <= 0) n = 1;
    cache.emplace_back(kern, n);
    return n;
}

static void demod_kernels(const Geom &g, const FrameJob *jobs, int n_jobs, const float *sine, const float *cosine,
                          const float *cfilt, int n_lblocks, hipStream_t s, int nchroma, float *raw)
{
    const bool composite = nchroma == 0;
    const int n_segs = (g.aw + kDemodThreads - 1) / kDemodThreads;
    if (g.quad && g.aw <= kQ4MaxWidth) {
        // A few frames (one chd_decode_frame window) give too few 32-line
        // workgroups to fill the chip, and each runs for the whole launch:
        // 8-line workgroups cut that latency at the cost of more halo rows.
        const int field_lines = n_lblocks * kDemodLines;   // at least every field's active lines
        const bool short_blocks = (int64_t)n_jobs * 2 * ((field_lines + kQ4Lines - 1) / kQ4Lines) < 256;
#define CHD_Q4(L)                                                                                              \
    (raw != nullptr ? (nchroma == 4 ? pal_demod_quad4<4, true, L> : nchroma == 2 ? pal_demod_quad4<2, true, L>     \
                       : nchroma == 1 ? pal_demod_quad4<1, true, L> : pal_demod_quad4<0, true, L>)               \
                    : (nchroma == 4 ? pal_demod_quad4<4, false, L> : nchroma == 2 ? pal_demod_quad4<2, false, L>   \
                       : nchroma == 1 ? pal_demod_quad4<1, false, L> : pal_demod_quad4<0, false, L>))
        auto kern = short_blocks ? CHD_Q4(kQ4ShortLines) : CHD_Q4(kQ4Lines);
#undef CHD_Q4
        // Every workgroup runs about as long as its lines plus the 6-line
        // halo, and the grid drains in rounds of the resident slots: pick
        // the lines per workgroup (16 .. 48) that makes the fewest rounds of
        // the shortest workgroups.  (For 64 PAL frames, 48 lines give 768
        // workgroups, one full round of 3 per CU; 24 give two rounds with
        // twice the halo rows, 32 one and a half.)
        int lpb = short_blocks ? kQ4ShortLines : kQ4Lines;
        if (!short_blocks) {
            const int64_t slots = (int64_t)std::max(1, resident_per_cu((const void *)kern)) * cu_count();
            int64_t best = -1;
            for (int cand = 16; cand <= kQ4Lines; cand += 4) {
                const int64_t wgs = (int64_t)n_jobs * 2 * ((field_lines + cand - 1) / cand);
                const int64_t cost = ((wgs + slots - 1) / slots) * (cand + 6);
                if (best < 0 || cost < best) {
                    best = cost;
                    lpb = cand;
                }
            }
        }
        const int nlb = (field_lines + lpb - 1) / lpb;
        const int blocks = n_jobs * 2 * nlb;
        if (blocks > 0)
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(kDemodThreads), 0, s, g, jobs, sine, cosine, cfilt, nlb, lpb,
                               raw);
        return;
    }
    if (g.quad && !composite) {  // active lines wider than one column-blocked workgroup
        const int nlb = (n_lblocks * kDemodLines + kQLines - 1) / kQLines;
        const int blocks = n_jobs * 2 * nlb * n_segs;
        if (blocks > 0)
            hipLaunchKernelGGL(pal_demod_quad, dim3(blocks), dim3(kDemodThreads), 0, s, g, jobs, sine, cosine, cfilt,
                               nlb, n_segs, raw);
        return;
    }
    const int blocks = n_jobs * 2 * n_lblocks * n_segs;
    if (blocks > 0)
        hipLaunchKernelGGL(pal_demod, dim3(blocks), dim3(kDemodThreads), 0, s, g, jobs, sine, cosine, cfilt,
                           n_lblocks, n_segs, raw);
}

hipError_t launch_demod(const Geom &g, const FrameJob *jobs, int n_jobs, const float *sine, const float *cosine,
                        const float *cfilt, int n_lblocks, hipStream_t s, int nchroma, float *raw)
{
    if (g.ynr > 0.0f && raw == nullptr) return hipErrorInvalidValue;  // coring reads the raw planes
    if (g.ynr <= 0.0f) raw = nullptr;
    demod_kernels(g, jobs, n_jobs, sine, cosine, cfilt, n_lblocks, s, nchroma, raw);
    if (raw != nullptr && n_jobs > 0)
        hipLaunchKernelGGL(ynr_output, dim3(n_jobs * (g.last - g.first + 1)), dim3(256), 0, s, g, jobs, raw);
    return hipSuccess;
}

void launch_pack(const Geom &g, const uint16_t *src, int64_t src_field_elems, uint16_t *dst, int n_fields,
                 hipStream_t s, int shift)
{
    if (n_fields > 0)
        hipLaunchKernelGGL(pack_fields, dim3(1024), dim3(256), 0, s, g, src, src_field_elems, dst, n_fields, shift);
}

void launch_copy_in(const void *src, void *dst, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return;
    const int64_t n16 = (int64_t)(bytes / 16);
    const int tail = (int)(bytes % 16);
    hipLaunchKernelGGL(copy_in, dim3(512), dim3(256), 0, s, static_cast<const uint4 *>(src), static_cast<uint4 *>(dst),
                       n16, static_cast<const uint8_t *>(src) + n16 * 16, static_cast<uint8_t *>(dst) + n16 * 16, tail);
}

void launch_pack_list(const Geom &g, const uint16_t *src, int64_t src_field_elems, uint16_t *dst,
                      const PackList &list, hipStream_t s)
{
    if (list.n > 0 && list.n <= kPackListMax)
        hipLaunchKernelGGL(pack_field_list, dim3(256, list.n), dim3(256), 0, s, g, src, src_field_elems, dst, list);
}

void launch_chroma_to_planes(const Geom &g, const float *src, float *dst, int64_t plane_stride, int n_fields,
                             hipStream_t s)
{
    if (n_fields > 0) hipLaunchKernelGGL(chroma_to_planes, dim3(1024), dim3(256), 0, s, g, src, dst, plane_stride, n_fields);
}

void launch_fill_border(const Geom &g, void *out, int64_t n_frames, hipStream_t s)
{
    if (n_frames > 0) hipLaunchKernelGGL(fill_border, dim3(1024), dim3(256), 0, s, g, out, n_frames);
}

void launch_fill_other_parity(const Geom &g, void *out, int parity, hipStream_t s)
{
    hipLaunchKernelGGL(fill_other_parity, dim3(512), dim3(256), 0, s, g, out, parity & 1);
}

}  // namespace tpal

#ifdef CHD_TILE_PHASES
// Copy the tile-phase sums (32 counters) to `out`; clear them when `clear`.
extern "C" __attribute__((visibility("default"))) int chd_tile_phases_read(unsigned long long *out, int clear)
{
    if (hipDeviceSynchronize() != hipSuccess) return -1;
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(tpal::g_tile_phases), sizeof(tpal::g_tile_phases)) != hipSuccess)
        return -1;
    if (clear) {
        static const unsigned long long zero[32] = {};
        if (hipMemcpyToSymbol(HIP_SYMBOL(tpal::g_tile_phases), zero, sizeof(zero)) != hipSuccess) return -1;
    }
    return 0;
}
#endif
