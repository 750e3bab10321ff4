/*
 * chd_oracle.c -- FP64 CPU restatement of libchromadec's Transform-PAL decode
 * path (CHD_DEC_TRANSFORM_2D / CHD_DEC_TRANSFORM_3D).
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  This file is the checker that the
 * HIP decoder is compared against.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py load it; the product library never links,
 * loads or calls it.
 *
 * It restates the reference algorithm step by step, in double precision and
 * with the same loop structure (full 3D/2D real-to-complex FFT, the
 * carrier-mirror threshold, complex-to-real inverse, overlap-add), so that it
 * checks the pruned FP32 formulation the GPU kernels use rather than sharing
 * it.  FFTW is not available here, so the transforms are a textbook radix-2
 * FFT (pinned against a brute-force DFT by tests/test_oracle_fft.py).
 *
 * Reference files followed (paths relative to the libchromadec tree):
 *   src/decoders/palcolour/transform_pal.cpp:53-77      thresholds (squared)
 *   src/decoders/palcolour/transform_pal_3d.cpp:64-333  3D window/lattice/filter/OLA
 *   src/decoders/palcolour/transform_pal_2d.cpp:63-287  2D window/lattice/filter/OLA
 *   src/decoders/palcolour/palcolour.cpp:150-409,456-646 LUTs, burst, demod
 *   src/decoders/palcolour/palcolour.cpp:411-448        luma noise coring (doYNR)
 *   src/decoders/filter/deemp.h:620-633, iirfilter.h:39-98  its high-pass FIR
 *   src/decoders/source_field.h:77-91                   field offset / active lines
 *   src/output/component_frame.cpp:35-61                component frame layout
 *   src/output/output_writer.cpp:46-76,124-143,371-486,670-727  output conversion
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ORC_MAX_WIDTH 1135 /* PalColour::MAX_WIDTH, palcolour.h:93 */

typedef struct orc_params {
    int32_t field_width, field_height;
    int32_t active_video_start, active_video_end;          /* half-open samples */
    int32_t first_active_frame_line, last_active_frame_line; /* inclusive frame lines */
    int32_t colour_burst_start, colour_burst_end;
    int32_t black16b, white16b;
    double fsc, sample_rate;
    double chroma_gain, chroma_phase_deg;
    double ecb_from_u, ecr_from_v;
    int32_t clamp_mode; /* 0 none, 1 legal_rgb_sdr, 2 legal_rgb_hdr, 3 legal_ycbcr_bt601 */
    int32_t pad_left, pad_top, out_width, out_height;
    double kr, kb;      /* luma coefficients of the colour-difference precision */
    double luma_nr_level; /* PalColour yNRLevel in IRE; 0 = off */
} orc_params;

/* ------------------------------------------------------------------------ */
/* Complex FFT (radix-2, iterative, unnormalised).  sign = -1 forward.       */
/* ------------------------------------------------------------------------ */

/* Twiddle tables for n = 2..64, filled once at load. */
static double tw_re[7][32], tw_im[7][32];

__attribute__((constructor)) static void init_twiddles(void)
{
    for (int lg = 1; lg <= 6; lg++) {
        const int len = 1 << lg;
        for (int k = 0; k < len / 2; k++) {
            tw_re[lg][k] = cos(2.0 * M_PI * k / len);
            tw_im[lg][k] = sin(2.0 * M_PI * k / len);
        }
    }
}

static void fft1d(double *re, double *im, int n, int stride, int sign)
{
    /* bit reversal */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = re[i * stride], ti = im[i * stride];
            re[i * stride] = re[j * stride]; im[i * stride] = im[j * stride];
            re[j * stride] = tr; im[j * stride] = ti;
        }
    }
    for (int lg = 1, len = 2; len <= n; len <<= 1, lg++) {
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < len / 2; k++) {
                const double wr = tw_re[lg][k], wi = sign * tw_im[lg][k];
                const int a = (i + k) * stride, b = (i + k + len / 2) * stride;
                const double xr = re[b] * wr - im[b] * wi;
                const double xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
        }
    }
}

/* In-place complex 3D FFT over a [Z][Y][X] array. */
void orc_fft3d(double *re, double *im, int Z, int Y, int X, int sign)
{
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            fft1d(re + (z * Y + y) * X, im + (z * Y + y) * X, X, 1, sign);
    for (int z = 0; z < Z; z++)
        for (int x = 0; x < X; x++)
            fft1d(re + z * Y * X + x, im + z * Y * X + x, Y, X, sign);
    for (int y = 0; y < Y; y++)
        for (int x = 0; x < X; x++)
            fft1d(re + y * X + x, im + y * X + x, Z, Y * X, sign);
}

/* Brute-force 3D DFT (forward, unnormalised) of a real [Z][Y][X] array.
 * Used only to pin orc_fft3d / the r2c wrapper in tests. */
void orc_dft3d_naive(const double *in, int Z, int Y, int X, double *out_re, double *out_im)
{
    for (int kz = 0; kz < Z; kz++)
        for (int ky = 0; ky < Y; ky++)
            for (int kx = 0; kx < X; kx++) {
                double sr = 0, si = 0;
                for (int z = 0; z < Z; z++)
                    for (int y = 0; y < Y; y++)
                        for (int x = 0; x < X; x++) {
                            const double ph = -2.0 * M_PI *
                                ((double)(kz * z) / Z + (double)(ky * y) / Y + (double)(kx * x) / X);
                            const double v = in[(z * Y + y) * X + x];
                            sr += v * cos(ph);
                            si += v * sin(ph);
                        }
                out_re[(kz * Y + ky) * X + kx] = sr;
                out_im[(kz * Y + ky) * X + kx] = si;
            }
}

/* Real-to-complex 3D FFT with FFTW's half-spectrum layout [Z][Y][X/2+1]
 * (fftw_plan_dft_r2c_3d semantics).  Scratch re/im hold Z*Y*X doubles. */
void orc_r2c3d(const double *in, int Z, int Y, int X, double *h_re, double *h_im,
               double *re, double *im)
{
    const int XC = X / 2 + 1;
    for (int i = 0; i < Z * Y * X; i++) { re[i] = in[i]; im[i] = 0.0; }
    /* x pass on every row; the y and z passes only need bins 0..X/2. */
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            fft1d(re + (z * Y + y) * X, im + (z * Y + y) * X, X, 1, -1);
    for (int z = 0; z < Z; z++)
        for (int x = 0; x < XC; x++)
            fft1d(re + z * Y * X + x, im + z * Y * X + x, Y, X, -1);
    for (int y = 0; y < Y; y++)
        for (int x = 0; x < XC; x++)
            fft1d(re + y * X + x, im + y * X + x, Z, Y * X, -1);
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            for (int x = 0; x < XC; x++) {
                h_re[(z * Y + y) * XC + x] = re[(z * Y + y) * X + x];
                h_im[(z * Y + y) * XC + x] = im[(z * Y + y) * X + x];
            }
}

/* Complex-to-real 3D inverse (fftw_plan_dft_c2r_3d semantics, unnormalised):
 * the half spectrum is completed by Hermitian symmetry and inverted. */
void orc_c2r3d(const double *h_re, const double *h_im, int Z, int Y, int X, double *out,
               double *re, double *im)
{
    const int XC = X / 2 + 1;
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            for (int x = 0; x < XC; x++) {
                re[(z * Y + y) * X + x] = h_re[(z * Y + y) * XC + x];
                im[(z * Y + y) * X + x] = h_im[(z * Y + y) * XC + x];
            }
    /* z and y inverse passes on bins 0..X/2; bins above follow from Hermitian
     * symmetry: G(z, y, X-kx) = conj(G(z, y, kx)) after those passes. */
    for (int y = 0; y < Y; y++)
        for (int x = 0; x < XC; x++)
            fft1d(re + y * X + x, im + y * X + x, Z, Y * X, +1);
    for (int z = 0; z < Z; z++)
        for (int x = 0; x < XC; x++)
            fft1d(re + z * Y * X + x, im + z * Y * X + x, Y, X, +1);
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++) {
            double *r = re + (z * Y + y) * X, *i = im + (z * Y + y) * X;
            for (int x = XC; x < X; x++) { r[x] = r[X - x]; i[x] = -i[X - x]; }
            fft1d(r, i, X, 1, +1);
        }
    for (int k = 0; k < Z * Y * X; k++) out[k] = re[k];
}

/* ------------------------------------------------------------------------ */
/* Window (transform_pal_3d.cpp:64-67, transform_pal_2d.cpp:63-66)           */
/* ------------------------------------------------------------------------ */

double orc_window(int element, int limit)
{
    return 0.5 - (0.5 * cos((2 * M_PI * (element + 0.5)) / limit));
}

/* ------------------------------------------------------------------------ */
/* Carrier-mirror filter, generic over tile size (transform_pal_3d.cpp:249-333,
 * transform_pal_2d.cpp:~200-275).  Spectrum layout [Z][Y][XC].  thr_sq has
 * (XC/4+1)*Y*Z entries in z, y, x loop order.  out_* is cleared first.
 * zt is ZTILE (1 for 2D: z_ref = 0 then).  yref_shift is YTILE/4 for 3D and
 * YTILE/2 for 2D; zref_shift is ZTILE/4.                                    */
/* ------------------------------------------------------------------------ */

static void apply_filter(const double *in_re, const double *in_im, double *out_re, double *out_im,
                         int ZT, int YT, int XT, int zref_shift, int yref_shift, const double *thr_sq)
{
    const int XC = XT / 2 + 1;
    memset(out_re, 0, sizeof(double) * ZT * YT * XC);
    memset(out_im, 0, sizeof(double) * ZT * YT * XC);
    const double *t = thr_sq;
    for (int z = 0; z < ZT; z++) {
        const int z_ref = (zref_shift + ZT - z) % ZT;
        for (int y = 0; y < YT; y++) {
            const int y_ref = (yref_shift + YT - y) % YT;
            const int bi = (z * YT + y) * XC;
            const int bi_ref = (z_ref * YT + y_ref) * XC;
            for (int x = XT / 8; x <= XT / 4; x++) {
                const int x_ref = (XT / 2) - x;
                const double threshold_sq = *t++;
                const double ar = in_re[bi + x], ai = in_im[bi + x];
                const double br = in_re[bi_ref + x_ref], bim = in_im[bi_ref + x_ref];
                if (x == x_ref && y == y_ref && z == z_ref) {
                    out_re[bi + x] = ar; out_im[bi + x] = ai;
                    continue;
                }
                const double m_in_sq = ar * ar + ai * ai;
                const double m_ref_sq = br * br + bim * bim;
                if (m_in_sq < m_ref_sq * threshold_sq || m_ref_sq < m_in_sq * threshold_sq) {
                    /* discard */
                } else {
                    out_re[bi + x] = ar; out_im[bi + x] = ai;
                    out_re[bi_ref + x_ref] = br; out_im[bi_ref + x_ref] = bim;
                }
            }
        }
    }
}

/* Exposed for tests: run the 3D filter on a given half spectrum. */
void orc_apply_filter3d(const double *in_re, const double *in_im, double *out_re, double *out_im,
                        const double *thr_sq)
{
    apply_filter(in_re, in_im, out_re, out_im, 8, 32, 16, 8 / 4, 32 / 4, thr_sq);
}

void orc_apply_filter2d(const double *in_re, const double *in_im, double *out_re, double *out_im,
                        const double *thr_sq)
{
    apply_filter(in_re, in_im, out_re, out_im, 1, 16, 32, 0, 16 / 2, thr_sq);
}

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* Sample read with a black fallback outside the stored row.  The reference
 * reads b[tileX + x] unchecked; with the default crop every read is inside
 * the row, and this guard only changes reads the reference leaves undefined. */
static inline double sample_at(const uint16_t *row, int x, int fw, int black)
{
    return (x >= 0 && x < fw) ? (double)row[x] : (double)black;
}

/* ------------------------------------------------------------------------ */
/* TransformPal3D::filterFields (transform_pal_3d.cpp:123-240)               */
/* fields[nfields]: field samples; offsets[i]: 0 first field, 1 second.       */
/* out[end-start]: chroma buffers of field_width*field_height doubles.       */
/* ------------------------------------------------------------------------ */

enum { ZT3 = 8, YT3 = 32, XT3 = 16, XC3 = XT3 / 2 + 1 };

/* Soft mask in place of applyFilter (the learned mask of this library; no
 * reference counterpart): bins x = 2..4 take their visit's mask value m[v],
 * v = (z*32 + y)*3 + (x-2); bins x = 5, 6 take the value of the visit whose
 * carrier-mirror partner they are, (z_ref, y_ref, 8-x); other bins are 0. */
void orc_apply_soft_mask3d(const double *in_re, const double *in_im, double *out_re, double *out_im,
                              const float *m)
{
    memset(out_re, 0, sizeof(double) * ZT3 * YT3 * XC3);
    memset(out_im, 0, sizeof(double) * ZT3 * YT3 * XC3);
    for (int z = 0; z < ZT3; z++) {
        const int z_ref = (2 + ZT3 - z) % ZT3;
        for (int y = 0; y < YT3; y++) {
            const int y_ref = (8 + YT3 - y) % YT3;
            const int bi = (z * YT3 + y) * XC3;
            for (int x = 2; x <= 6; x++) {
                const double mv = x <= 4 ? m[(z * YT3 + y) * 3 + (x - 2)] : m[(z_ref * YT3 + y_ref) * 3 + (6 - x)];
                out_re[bi + x] = in_re[bi + x] * mv;
                out_im[bi + x] = in_im[bi + x] * mv;
            }
        }
    }
}

/* masks == NULL: the reference's threshold rule.  Otherwise masks holds 768
 * values per tile, tiles in (tileZ, tileY, tileX) loop order. */
static void tpal3d_filter(const orc_params *p, const uint16_t *const *fields, const int32_t *offsets,
                          int32_t start, int32_t end, const double *thr_sq, const float *masks, double *const *out)
{
    const int fw = p->field_width, fh = p->field_height;
    for (int i = 0; i < end - start; i++) memset(out[i], 0, sizeof(double) * fw * fh);

    double (*win)[YT3][XT3] = malloc(sizeof(double) * ZT3 * YT3 * XT3);
    for (int z = 0; z < ZT3; z++)
        for (int y = 0; y < YT3; y++)
            for (int x = 0; x < XT3; x++)
                win[z][y][x] = orc_window(z, ZT3) * orc_window(y, YT3) * orc_window(x, XT3);

    double *real = malloc(sizeof(double) * ZT3 * YT3 * XT3);
    double *sre = malloc(sizeof(double) * ZT3 * YT3 * XT3);
    double *sim = malloc(sizeof(double) * ZT3 * YT3 * XT3);
    double *hin_re = malloc(sizeof(double) * ZT3 * YT3 * XC3);
    double *hin_im = malloc(sizeof(double) * ZT3 * YT3 * XC3);
    double *hout_re = malloc(sizeof(double) * ZT3 * YT3 * XC3);
    double *hout_im = malloc(sizeof(double) * ZT3 * YT3 * XC3);

    const int first = p->first_active_frame_line, last = p->last_active_frame_line;
    const int avs = p->active_video_start, ave = p->active_video_end;
    const double black = p->black16b;

    for (int tileZ = start - ZT3 / 2; tileZ < end; tileZ += ZT3 / 2) {
        for (int tileY = first - YT3 / 2; tileY <= last; tileY += YT3 / 2) {
            for (int tileX = avs - XT3 / 2; tileX < ave; tileX += XT3 / 2) {
                /* forwardFFTTile (:168-202) */
                const int startY = imax(first - tileY, 0);
                const int endY = imin((last - tileY) + 1, YT3);
                for (int z = 0; z < ZT3; z++) {
                    const int fi = tileZ + z;
                    const int off = offsets[fi];
                    const uint16_t *in = fields[fi];
                    for (int y = 0; y < YT3; y++) {
                        double *dst = real + (z * YT3 + y) * XT3;
                        if (y < startY || y >= endY || ((tileY + y) % 2) != off) {
                            for (int x = 0; x < XT3; x++) dst[x] = black * win[z][y][x];
                            continue;
                        }
                        const uint16_t *b = in + ((tileY + y) / 2) * fw;
                        for (int x = 0; x < XT3; x++)
                            dst[x] = sample_at(b, tileX + x, fw, p->black16b) * win[z][y][x];
                    }
                }
                orc_r2c3d(real, ZT3, YT3, XT3, hin_re, hin_im, sre, sim);

                /* applyFilter (:249-333) */
                if (masks != NULL) {
                    orc_apply_soft_mask3d(hin_re, hin_im, hout_re, hout_im, masks);
                    masks += 768;
                } else {
                    apply_filter(hin_re, hin_im, hout_re, hout_im, ZT3, YT3, XT3, ZT3 / 4, YT3 / 4, thr_sq);
                }

                /* inverseFFTTile (:205-240) */
                const int startX = imax(avs - tileX, 0);
                const int endX = imin(ave - tileX, XT3);
                const int startZ = imax(start - tileZ, 0);
                const int endZ = imin(end - tileZ, ZT3);
                orc_c2r3d(hout_re, hout_im, ZT3, YT3, XT3, real, sre, sim);
                for (int z = startZ; z < endZ; z++) {
                    const int fi = tileZ + z;
                    const int oi = fi - start;
                    const int off = offsets[fi];
                    double *o = out[oi];
                    for (int y = startY; y < endY; y++) {
                        if (((tileY + y) % 2) != off) continue;
                        double *b = o + ((tileY + y) / 2) * fw;
                        for (int x = startX; x < endX; x++)
                            b[tileX + x] += real[(z * YT3 + y) * XT3 + x] / (ZT3 * YT3 * XT3);
                    }
                }
            }
        }
    }
    free(win);
    free(real); free(sre); free(sim);
    free(hin_re); free(hin_im); free(hout_re); free(hout_im);
}

void orc_tpal3d_filter_fields(const orc_params *p, const uint16_t *const *fields, const int32_t *offsets,
                              int32_t nfields, int32_t start, int32_t end, const double *thr_sq,
                              double *const *out)
{
    (void)nfields;
    tpal3d_filter(p, fields, offsets, start, end, thr_sq, NULL, out);
}

/* ------------------------------------------------------------------------ */
/* TransformPal2D::filterFields / filterField (transform_pal_2d.cpp:107-220)  */
/* ------------------------------------------------------------------------ */

enum { YT2 = 16, XT2 = 32, XC2 = XT2 / 2 + 1 };

void orc_tpal2d_filter_fields(const orc_params *p, const uint16_t *const *fields, const int32_t *offsets,
                              int32_t nfields, int32_t start, int32_t end, const double *thr_sq,
                              double *const *out)
{
    (void)nfields;
    const int fw = p->field_width, fh = p->field_height;
    double win[YT2][XT2];
    for (int y = 0; y < YT2; y++)
        for (int x = 0; x < XT2; x++) win[y][x] = orc_window(y, YT2) * orc_window(x, XT2);

    double real[YT2 * XT2], sre[YT2 * XT2], sim[YT2 * XT2];
    double hin_re[YT2 * XC2], hin_im[YT2 * XC2], hout_re[YT2 * XC2], hout_im[YT2 * XC2];
    const int avs = p->active_video_start, ave = p->active_video_end;

    for (int i = start, j = 0; i < end; i++, j++) {
        double *o = out[j];
        memset(o, 0, sizeof(double) * fw * fh);
        const int off = offsets[i];
        const uint16_t *in = fields[i];
        /* SourceField::getFirstActiveLine / getLastActiveLine (source_field.h:86-91) */
        const int firstFieldLine = (p->first_active_frame_line + 1 - off) / 2;
        const int lastFieldLine = ((p->last_active_frame_line - off) / 2) + 1;
        for (int tileY = firstFieldLine - YT2 / 2; tileY < lastFieldLine; tileY += YT2 / 2) {
            const int startY = imax(firstFieldLine - tileY, 0);
            const int endY = imin(lastFieldLine - tileY, YT2);
            for (int tileX = avs - XT2 / 2; tileX < ave; tileX += XT2 / 2) {
                for (int y = 0; y < YT2; y++) {
                    double *dst = real + y * XT2;
                    if (y < startY || y >= endY) {
                        for (int x = 0; x < XT2; x++) dst[x] = p->black16b * win[y][x];
                        continue;
                    }
                    const uint16_t *b = in + (tileY + y) * fw;
                    for (int x = 0; x < XT2; x++)
                        dst[x] = sample_at(b, tileX + x, fw, p->black16b) * win[y][x];
                }
                orc_r2c3d(real, 1, YT2, XT2, hin_re, hin_im, sre, sim);
                apply_filter(hin_re, hin_im, hout_re, hout_im, 1, YT2, XT2, 0, YT2 / 2, thr_sq);
                const int startX = imax(avs - tileX, 0);
                const int endX = imin(ave - tileX, XT2);
                orc_c2r3d(hout_re, hout_im, 1, YT2, XT2, real, sre, sim);
                for (int y = startY; y < endY; y++) {
                    double *b = o + (tileY + y) * fw;
                    for (int x = startX; x < endX; x++) b[tileX + x] += real[y * XT2 + x] / (YT2 * XT2);
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* PalColour (palcolour.cpp:150-409, 456-646), PREFILTERED_CHROMA branch      */
/* ------------------------------------------------------------------------ */

typedef struct {
    double sine[ORC_MAX_WIDTH], cosine[ORC_MAX_WIDTH];
    int filter_size;
    double cfilt[64][4];
    double yfilt[64][2];
} pal_luts;

static double dmin(double a, double b) { return a < b ? a : b; }

/* buildLookUpTables (palcolour.cpp:150-277) */
static void build_luts(const orc_params *p, pal_luts *L)
{
    for (int i = 0; i < p->field_width && i < ORC_MAX_WIDTH; i++) {
        const double rad = 2 * M_PI * i * p->fsc / p->sample_rate;
        if (p->field_height != 263) {
            L->sine[i] = sin(rad); L->cosine[i] = cos(rad);
        } else {
            L->sine[i] = cos(rad); L->cosine[i] = sin(rad);
        }
    }
    const double chromaBandwidthHz = 1100000.0 / 0.93;
    const double ca = 0.5 * p->sample_rate / chromaBandwidthHz;
    const double ya = 0.5 * p->sample_rate / chromaBandwidthHz;
    int fs = 1;
    if ((int)ca > fs) fs = (int)ca;
    if ((int)ya > fs) fs = (int)ya;
    if (fs > 63) fs = 63;
    L->filter_size = fs;
    double cdiv = 0, ydiv = 0;
    for (int f = 0; f <= fs; f++) {
        const double fc = dmin(ca, (double)f);
        const double ff = dmin(ca, sqrt(f * f + 2 * 2));
        const double fff = dmin(ca, sqrt(f * f + 4 * 4));
        const double ffff = dmin(ca, sqrt(f * f + 6 * 6));
        const int d = (f == 0) ? 2 : 1;
        L->cfilt[f][0] = (1 + cos(M_PI * fc / ca)) / d;
        L->cfilt[f][2] = (1 + cos(M_PI * ff / ca)) / d;
        L->cfilt[f][1] = (1 + cos(M_PI * fff / ca)) / d;
        L->cfilt[f][3] = (1 + cos(M_PI * ffff / ca)) / d;
        cdiv += 2 * (1 * L->cfilt[f][0] + 2 * L->cfilt[f][2] + 2 * L->cfilt[f][1] + 2 * L->cfilt[f][3]);
        const double fy = dmin(ya, (double)f);
        const double fffy = dmin(ya, sqrt(f * f + 4 * 4));
        L->yfilt[f][0] = (1 + cos(M_PI * fy / ya)) / d;
        L->yfilt[f][1] = 0.2 * (1 + cos(M_PI * fffy / ya)) / d;
        ydiv += 2 * (1 * L->yfilt[f][0] + 2 * 0 + 2 * L->yfilt[f][1] + 2 * 0);
    }
    for (int f = 0; f <= fs; f++) {
        for (int i = 0; i < 4; i++) L->cfilt[f][i] /= cdiv;
        for (int i = 0; i < 2; i++) L->yfilt[f][i] /= ydiv;
    }
}

/* Exposed for tests: the demodulator tables. */
void orc_pal_luts(const orc_params *p, double *sine, double *cosine, int32_t *filter_size,
                  double *cfilt /* [64][4] */, double *yfilt /* [64][2] */)
{
    pal_luts *L = malloc(sizeof(pal_luts));
    build_luts(p, L);
    memcpy(sine, L->sine, sizeof(double) * ORC_MAX_WIDTH);
    memcpy(cosine, L->cosine, sizeof(double) * ORC_MAX_WIDTH);
    *filter_size = L->filter_size;
    memcpy(cfilt, L->cfilt, sizeof(L->cfilt));
    memcpy(yfilt, L->yfilt, sizeof(L->yfilt));
    free(L);
}

/* detectBurst (palcolour.cpp:344-409).  Writes bp, bq, Vsw. */
static void detect_burst(const orc_params *p, const pal_luts *L, int number, const uint16_t *in,
                         double *obp, double *obq, double *oVsw)
{
    static const uint16_t blackLine[ORC_MAX_WIDTH] = {0};
    const int fw = p->field_width, fh = p->field_height;
    const uint16_t *in0 = in + number * fw;
    const uint16_t *in1 = (number - 1) < 0 ? blackLine : in + (number - 1) * fw;
    const uint16_t *in2 = (number + 1) >= fh ? blackLine : in + (number + 1) * fw;
    const uint16_t *in3 = (number - 2) < 0 ? blackLine : in + (number - 2) * fw;
    const uint16_t *in4 = (number + 2) >= fh ? blackLine : in + (number + 2) * fw;
    double bp = 0, bq = 0, bpo = 0, bqo = 0;
    for (int i = p->colour_burst_start; i < p->colour_burst_end; i++) {
        bp += ((in0[i] - ((in3[i] + in4[i]) / 2.0)) / 2.0) * L->sine[i];
        bq += ((in0[i] - ((in3[i] + in4[i]) / 2.0)) / 2.0) * L->cosine[i];
        bpo += ((in2[i] - in1[i]) / 2.0) * L->sine[i];
        bqo += ((in2[i] - in1[i]) / 2.0) * L->cosine[i];
    }
    const int len = p->colour_burst_end - p->colour_burst_start;
    bp /= len; bq /= len; bpo /= len; bqo /= len;
    double Vsw = -1;
    if (((bp - bpo) * (bp - bpo) + (bq - bqo) * (bq - bqo)) < (bp * bp + bq * bq) * 2) Vsw = 1;
    double lbp = (bp - bqo) / 2;
    double lbq = (bq + bpo) / 2;
    double norm = sqrt(lbp * lbp + lbq * lbq);
    if (norm < 130000.0 / 128) norm = 130000.0 / 128;
    *obp = lbp / norm;
    *obq = lbq / norm;
    *oVsw = Vsw;
}

/* Exposed for tests. */
void orc_detect_burst(const orc_params *p, int32_t number, const uint16_t *field,
                      double *bp, double *bq, double *vsw)
{
    pal_luts *L = malloc(sizeof(pal_luts));
    build_luts(p, L);
    detect_burst(p, L, number, field, bp, bq, vsw);
    free(L);
}

/* decodeField + decodeLine<double, true> (palcolour.cpp:307-335, 456-646).
 * Y/U/V are ComponentFrame planes of field_width x (2*field_height-1).
 * With pal2d set, `chroma` holds the composite field itself (as doubles) and
 * the line is decoded as decodeLine<uint16_t, false> (the PALcolour 2D
 * filter on the raw composite, palcolour.cpp:327-328, 525-609, 622-630):
 * luma subtracts the chroma re-synthesised from the Y filter's P/Q. */
/* PalColour::doYNR (palcolour.cpp:411-448): analog-style luma coring of one
 * frame line.  f_nrpal is an IIRFilter<25, 1> (deemp.h:620-633), a 25-tap
 * FIR with a[0] = 1; it is fed zeros for the 12 samples before the active
 * line, the active samples, then 12 zeros, and its output 12 samples later
 * is the high-pass of sample h.  That value, clipped to +-nr_y, is
 * subtracted.  nr_y = level * ((white - black) / 100), the division being
 * that of the int32 IRE fields. */
static const double c_nrpal_b[25] = {
    1.141291975113614e-04, -1.857019211291029e-03, -4.499636864042073e-03, -5.577680979937061e-03,
    -4.423694440267179e-04, 1.309163063177155e-02, 2.861211356202848e-02, 3.029931283148555e-02,
    1.098965697652802e-03, -6.398130386469833e-02, -1.492080690537196e-01, -2.223459379380252e-01,
    7.479077367478024e-01, -2.223459379380252e-01, -1.492080690537196e-01, -6.398130386469833e-02,
    1.098965697652803e-03, 3.029931283148557e-02, 2.861211356202848e-02, 1.309163063177156e-02,
    -4.423694440267185e-04, -5.577680979937061e-03, -4.499636864042074e-03, -1.857019211291030e-03,
    1.141291975113614e-04};

void orc_ynr_line(const orc_params *p, double *Yline)
{
    const double irescale = (double)((p->white16b - p->black16b) / 100);
    const double nr_y = p->luma_nr_level * irescale;
    const int delay = 25 / 2;
    const int avs = p->active_video_start, ave = p->active_video_end;
    double x[25] = {0};
    double *hp = calloc((size_t)ave + delay, sizeof(double));
    /* IIRFilter::feed with bOrder 25, aOrder 1 */
#define ORC_FEED(v, dst)                                  \
    do {                                                  \
        double y0 = c_nrpal_b[0] * (v);                   \
        for (int k = 24; k >= 1; k--) {                   \
            x[k] = x[k - 1];                              \
            y0 += c_nrpal_b[k] * x[k];                    \
        }                                                 \
        x[0] = (v);                                       \
        dst = y0;                                         \
    } while (0)
    double sink;
    for (int h = avs - delay; h < avs; h++) ORC_FEED(0.0, sink);
    for (int h = avs; h < ave; h++) ORC_FEED(Yline[h], hp[h]);
    for (int h = ave; h < ave + delay; h++) ORC_FEED(0.0, hp[h]);
#undef ORC_FEED
    (void)sink;
    for (int h = avs; h < ave; h++) {
        double a = hp[h + delay];
        if (fabs(a) > nr_y) a = (a > 0) ? nr_y : -nr_y;
        Yline[h] -= a;
    }
    free(hp);
}

static void decode_field(const orc_params *p, const pal_luts *L, const uint16_t *comp, int off,
                         const double *chroma, int pal2d, double *Yp, double *Up, double *Vp)
{
    static const double blackLine[ORC_MAX_WIDTH] = {0};
    const int fw = p->field_width;
    const int firstLine = (p->first_active_frame_line + 1 - off) / 2;
    const int lastLine = ((p->last_active_frame_line - off) / 2) + 1;
    const int fs = L->filter_size;
    const double theta = (p->chroma_phase_deg * M_PI) / 180;
    double (*m)[ORC_MAX_WIDTH] = malloc(sizeof(double) * 4 * ORC_MAX_WIDTH);
    double (*n)[ORC_MAX_WIDTH] = malloc(sizeof(double) * 4 * ORC_MAX_WIDTH);
    for (int number = firstLine; number < lastLine; number++) {
        double bp, bq, Vsw;
        detect_burst(p, L, number, comp, &bp, &bq, &Vsw);
        const double obp = bp, obq = bq;
        bp = (obp * cos(theta) - obq * sin(theta)) * p->chroma_gain;
        bq = (obp * sin(theta) + obq * cos(theta)) * p->chroma_gain;

        const double *in0 = chroma + number * fw;
        const double *in1 = (number - 1) < firstLine ? blackLine : chroma + (number - 1) * fw;
        const double *in2 = (number + 1) >= lastLine ? blackLine : chroma + (number + 1) * fw;
        const double *in3 = (number - 2) < firstLine ? blackLine : chroma + (number - 2) * fw;
        const double *in4 = (number + 2) >= lastLine ? blackLine : chroma + (number + 2) * fw;
        const double *in5 = (number - 3) < firstLine ? blackLine : chroma + (number - 3) * fw;
        const double *in6 = (number + 3) >= lastLine ? blackLine : chroma + (number + 3) * fw;

        const int endPos = imin(p->active_video_end, ORC_MAX_WIDTH);
        const int fillStart = imax(0, p->active_video_start - fs);
        const int fillEnd = imin(p->active_video_end + fs + 1, ORC_MAX_WIDTH);
        const double *S = L->sine, *C = L->cosine;
        for (int i = fillStart; i < fillEnd; i++) {
            m[0][i] = in0[i] * S[i];
            m[2][i] = in1[i] * S[i] - in2[i] * S[i];
            m[1][i] = -in3[i] * S[i] - in4[i] * S[i];
            m[3][i] = -in5[i] * S[i] + in6[i] * S[i];
            n[0][i] = in0[i] * C[i];
            n[2][i] = in1[i] * C[i] - in2[i] * C[i];
            n[1][i] = -in3[i] * C[i] - in4[i] * C[i];
            n[3][i] = -in5[i] * C[i] + in6[i] * C[i];
        }
        const uint16_t *cl = comp + number * fw;
        const int lineNumber = number * 2 + off;
        double *oY = Yp + lineNumber * fw, *oU = Up + lineNumber * fw, *oV = Vp + lineNumber * fw;
        for (int i = p->active_video_start; i < endPos; i++) {
            double PU = 0, QU = 0, PV = 0, QV = 0, PY = 0, QY = 0;
            for (int b = 0; b <= fs; b++) {
                const int l = imax(fillStart, i - b);
                const int r = imin(fillEnd - 1, i + b);
                PY += (m[0][r] + m[0][l]) * L->yfilt[b][0] + (m[1][r] + m[1][l]) * L->yfilt[b][1];
                QY += (n[0][r] + n[0][l]) * L->yfilt[b][0] + (n[1][r] + n[1][l]) * L->yfilt[b][1];
                PU += (m[0][r] + m[0][l]) * L->cfilt[b][0] + (m[1][r] + m[1][l]) * L->cfilt[b][1]
                    + (n[2][r] + n[2][l]) * L->cfilt[b][2] + (n[3][r] + n[3][l]) * L->cfilt[b][3];
                QU += (n[0][r] + n[0][l]) * L->cfilt[b][0] + (n[1][r] + n[1][l]) * L->cfilt[b][1]
                    - (m[2][r] + m[2][l]) * L->cfilt[b][2] - (m[3][r] + m[3][l]) * L->cfilt[b][3];
                PV += (m[0][r] + m[0][l]) * L->cfilt[b][0] + (m[1][r] + m[1][l]) * L->cfilt[b][1]
                    - (n[2][r] + n[2][l]) * L->cfilt[b][2] - (n[3][r] + n[3][l]) * L->cfilt[b][3];
                QV += (n[0][r] + n[0][l]) * L->cfilt[b][0] + (n[1][r] + n[1][l]) * L->cfilt[b][1]
                    + (m[2][r] + m[2][l]) * L->cfilt[b][2] + (m[3][r] + m[3][l]) * L->cfilt[b][3];
            }
            oY[i] = pal2d ? cl[i] - ((PY * S[i] + QY * C[i]) * 2.0) : cl[i] - in0[i];
            oU[i] = -(PU * bp + QU * bq) * 2.0;
            oV[i] = Vsw * -(QV * bp - PV * bq) * 2.0;
        }
        if (p->luma_nr_level > 0.0) orc_ynr_line(p, oY);
    }
    free(m);
    free(n);
}

/* PalColour::decodeFrames for one frame (two fields), given the chroma. */
void orc_palcolour_decode_frame(const orc_params *p, const uint16_t *field0, int32_t off0,
                                const uint16_t *field1, int32_t off1, const double *chroma0,
                                const double *chroma1, double *Y, double *U, double *V)
{
    const size_t n = (size_t)p->field_width * (2 * p->field_height - 1);
    memset(Y, 0, n * sizeof(double));
    memset(U, 0, n * sizeof(double));
    memset(V, 0, n * sizeof(double));
    pal_luts *L = malloc(sizeof(pal_luts));
    build_luts(p, L);
    decode_field(p, L, field0, off0, chroma0, 0, Y, U, V);
    decode_field(p, L, field1, off1, chroma1, 0, Y, U, V);
    free(L);
}

/* PalColour::decodeFrames with separation palColourFilter (CHD_DEC_PAL_2D):
 * each field decoded from its own composite (palcolour.cpp:279-335). */
void orc_pal2d_decode_frame(const orc_params *p, const uint16_t *field0, int32_t off0, const uint16_t *field1,
                            int32_t off1, double *Y, double *U, double *V)
{
    const size_t n = (size_t)p->field_width * (2 * p->field_height - 1);
    const size_t fa = (size_t)p->field_width * p->field_height;
    memset(Y, 0, n * sizeof(double));
    memset(U, 0, n * sizeof(double));
    memset(V, 0, n * sizeof(double));
    pal_luts *L = malloc(sizeof(pal_luts));
    double *c = malloc(2 * fa * sizeof(double));
    build_luts(p, L);
    for (size_t i = 0; i < fa; i++) { c[i] = field0[i]; c[fa + i] = field1[i]; }
    decode_field(p, L, field0, off0, c, 1, Y, U, V);
    decode_field(p, L, field1, off1, c + fa, 1, Y, U, V);
    free(c);
    free(L);
}

/* ------------------------------------------------------------------------ */
/* OutputWriter (output_writer.cpp:46-76, 124-143, 371-486, 670-727)          */
/* ------------------------------------------------------------------------ */

static double h273_round(double x) { return (x < 0.0 ? -1.0 : 1.0) * floor(fabs(x) + 0.5); }

static double dclamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
static float fclamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

/* YUV444PS: three float planes [Y|Cb|Cr] of out_width x out_height. */
void orc_output_float(const orc_params *p, const double *Y, const double *U, const double *V, float *out)
{
    const int W = p->out_width, H = p->out_height, fw = p->field_width;
    const size_t ps = (size_t)W * H;
    memset(out, 0, 3 * ps * sizeof(float));
    const double yOffset = p->black16b;
    const double yRange = p->white16b - p->black16b;
    const double eyScale = 1.0 / yRange;
    const double ecbScale = p->ecb_from_u / yRange;
    const double ecrScale = p->ecr_from_v / yRange;
    float eyLo = -INFINITY, eyHi = INFINITY, ecLo = -INFINITY, ecHi = INFINITY;
    if (p->clamp_mode == 1) { eyLo = 0.0f; eyHi = 1.0f; ecLo = -0.5f; ecHi = 0.5f; }
    if (p->clamp_mode == 3) {
        eyLo = -15.0f / 219.0f; eyHi = 238.75f / 219.0f;
        ecLo = -127.0f / 224.0f; ecHi = 126.75f / 224.0f;
    }
    const int aw = p->active_video_end - p->active_video_start;
    const int ah = p->last_active_frame_line - p->first_active_frame_line + 1;
    for (int y = 0; y < ah; y++) {
        const int il = p->first_active_frame_line + y;
        const size_t ro = (size_t)W * (p->pad_top + y) + p->pad_left;
        const double *iy = Y + (size_t)il * fw + p->active_video_start;
        const double *iu = U + (size_t)il * fw + p->active_video_start;
        const double *iv = V + (size_t)il * fw + p->active_video_start;
        for (int x = 0; x < aw; x++) {
            out[ro + x] = fclamp((float)((iy[x] - yOffset) * eyScale), eyLo, eyHi);
            out[ps + ro + x] = fclamp((float)(iu[x] * ecbScale), ecLo, ecHi);
            out[2 * ps + ro + x] = fclamp((float)(iv[x] * ecrScale), ecLo, ecHi);
        }
    }
}

/* YUV444P16: three u16 planes [Y|Cb|Cr] of out_width x out_height. */
void orc_output_u16(const orc_params *p, const double *Y, const double *U, const double *V, uint16_t *out)
{
    const int W = p->out_width, H = p->out_height, fw = p->field_width;
    const size_t ps = (size_t)W * H;
    for (size_t i = 0; i < ps; i++) { out[i] = 16 * 256; out[ps + i] = 128 * 256; out[2 * ps + i] = 128 * 256; }
    const double yOffset = p->black16b;
    const double yRange = p->white16b - p->black16b;
    const double eyScale = 1.0 / yRange;
    const double ecbScale = p->ecb_from_u / yRange;
    const double ecrScale = p->ecr_from_v / yRange;
    double yLo = 0, yHi = 65535, cLo = 0, cHi = 65535;
    if (p->clamp_mode == 1) { yLo = 16 * 256.0; yHi = 235 * 256.0; cLo = 16 * 256.0; cHi = 240 * 256.0; }
    if (p->clamp_mode == 3) { yLo = 256.0; yHi = 254.75 * 256.0; cLo = 256.0; cHi = 254.75 * 256.0; }
    const int aw = p->active_video_end - p->active_video_start;
    const int ah = p->last_active_frame_line - p->first_active_frame_line + 1;
    for (int y = 0; y < ah; y++) {
        const int il = p->first_active_frame_line + y;
        const size_t ro = (size_t)W * (p->pad_top + y) + p->pad_left;
        const double *iy = Y + (size_t)il * fw + p->active_video_start;
        const double *iu = U + (size_t)il * fw + p->active_video_start;
        const double *iv = V + (size_t)il * fw + p->active_video_start;
        for (int x = 0; x < aw; x++) {
            out[ro + x] = (uint16_t)dclamp(h273_round(219.0 * 256.0 * ((iy[x] - yOffset) * eyScale) + 16.0 * 256.0), yLo, yHi);
            out[ps + ro + x] = (uint16_t)dclamp(h273_round(224.0 * 256.0 * (iu[x] * ecbScale) + 128.0 * 256.0), cLo, cHi);
            out[2 * ps + ro + x] = (uint16_t)dclamp(h273_round(224.0 * 256.0 * (iv[x] * ecrScale) + 128.0 * 256.0), cLo, cHi);
        }
    }
}

/* R'G'B' from the composite-domain Y/U/V (color_conversion.cpp:33-67): the
 * composite U/V reductions follow from ecbFromU = 1 / (2 (1 - kb) uReduction)
 * and ecrFromV = 1 / (2 (1 - kr) vReduction). */
static void rgb_coeffs(const orc_params *p, double *rV, double *gU, double *gV, double *bU)
{
    const double kg = 1.0 - p->kr - p->kb;
    const double ured = 1.0 / (2.0 * (1.0 - p->kb) * p->ecb_from_u);
    const double vred = 1.0 / (2.0 * (1.0 - p->kr) * p->ecr_from_v);
    *rV = 1.0 / vred;
    *gU = -(p->kb / kg) / ured;
    *gV = -(p->kr / kg) / vred;
    *bU = 1.0 / ured;
}

/* RGB48: interleaved R'G'B' u16, full range, truncated after the [0, 65535]
 * clamp; black border (OutputWriter::convertLine, output_writer.cpp:431-452;
 * fillBlack :371-378). */
void orc_output_rgb48(const orc_params *p, const double *Y, const double *U, const double *V, uint16_t *out)
{
    const int W = p->out_width, H = p->out_height, fw = p->field_width;
    memset(out, 0, (size_t)3 * W * H * sizeof(uint16_t));
    double rV, gU, gV, bU;
    rgb_coeffs(p, &rV, &gU, &gV, &bU);
    const double yRange = p->white16b - p->black16b;
    const double yScale = 65535.0 / yRange, uvScale = 65535.0 / yRange;
    const int aw = p->active_video_end - p->active_video_start;
    const int ah = p->last_active_frame_line - p->first_active_frame_line + 1;
    for (int y = 0; y < ah; y++) {
        const int il = p->first_active_frame_line + y;
        const size_t ro = (size_t)W * (p->pad_top + y) + p->pad_left;
        const double *iy = Y + (size_t)il * fw + p->active_video_start;
        const double *iu = U + (size_t)il * fw + p->active_video_start;
        const double *iv = V + (size_t)il * fw + p->active_video_start;
        for (int x = 0; x < aw; x++) {
            const double ry = dclamp((iy[x] - p->black16b) * yScale, 0.0, 65535.0);
            const double ru = iu[x] * uvScale, rv = iv[x] * uvScale;
            uint16_t *o = out + (ro + x) * 3;
            o[0] = (uint16_t)dclamp(ry + rV * rv, 0.0, 65535.0);
            o[1] = (uint16_t)dclamp(ry + gU * ru + gV * rv, 0.0, 65535.0);
            o[2] = (uint16_t)dclamp(ry + bU * ru, 0.0, 65535.0);
        }
    }
}

/* RGBS: planar E'R/E'G/E'B floats, black 0, clamped to the mode's R'G'B'
 * box (OutputWriter::convertToFloatRGB, output_writer.cpp:729-777; bounds
 * floatRGBBounds :149-190, with the BT.601 Y'CbCr box projected through the
 * inverse matrix). */
void orc_output_rgbs(const orc_params *p, const double *Y, const double *U, const double *V, float *out)
{
    const int W = p->out_width, H = p->out_height, fw = p->field_width;
    const size_t ps = (size_t)W * H;
    memset(out, 0, 3 * ps * sizeof(float));
    double rV, gU, gV, bU;
    rgb_coeffs(p, &rV, &gU, &gV, &bU);
    float lo[3] = {-INFINITY, -INFINITY, -INFINITY}, hi[3] = {INFINITY, INFINITY, INFINITY};
    if (p->clamp_mode == 1) {
        for (int c = 0; c < 3; c++) { lo[c] = 0.0f; hi[c] = 1.0f; }
    } else if (p->clamp_mode == 2) {
        for (int c = 0; c < 3; c++) lo[c] = 0.0f;
    } else if (p->clamp_mode == 3) {
        const float eyLo = -15.0f / 219.0f, eyHi = 238.75f / 219.0f;
        const float ecLo = -127.0f / 224.0f, ecHi = 126.75f / 224.0f;
        const double kg = 1.0 - p->kr - p->kb;
        const double rds = 2.0 * (1.0 - p->kr), bds = 2.0 * (1.0 - p->kb);
        const double g = (p->kb * bds) / kg + (p->kr * rds) / kg;
        lo[0] = (float)(eyLo + rds * ecLo); hi[0] = (float)(eyHi + rds * ecHi);
        lo[1] = (float)(eyLo - g * ecHi);   hi[1] = (float)(eyHi - g * ecLo);
        lo[2] = (float)(eyLo + bds * ecLo); hi[2] = (float)(eyHi + bds * ecHi);
    }
    const double yRange = p->white16b - p->black16b;
    const double eyScale = 1.0 / yRange, uvScale = 1.0 / yRange;
    const int aw = p->active_video_end - p->active_video_start;
    const int ah = p->last_active_frame_line - p->first_active_frame_line + 1;
    for (int y = 0; y < ah; y++) {
        const int il = p->first_active_frame_line + y;
        const size_t ro = (size_t)W * (p->pad_top + y) + p->pad_left;
        const double *iy = Y + (size_t)il * fw + p->active_video_start;
        const double *iu = U + (size_t)il * fw + p->active_video_start;
        const double *iv = V + (size_t)il * fw + p->active_video_start;
        for (int x = 0; x < aw; x++) {
            const double ry = (iy[x] - p->black16b) * eyScale;
            const double ru = iu[x] * uvScale, rv = iv[x] * uvScale;
            out[ro + x] = fclamp((float)(ry + rV * rv), lo[0], hi[0]);
            out[ps + ro + x] = fclamp((float)(ry + gU * ru + gV * rv), lo[1], hi[1]);
            out[2 * ps + ro + x] = fclamp((float)(ry + bU * ru), lo[2], hi[2]);
        }
    }
}

/* out_fmt 0 = float Y'CbCr planes, 1 = u16 Y'CbCr planes, 2 = RGB48, 3 = RGBS */
static void write_output(const orc_params *p, const double *yuv, size_t ca, int32_t out_fmt, void *out)
{
    if (out == NULL) return;
    if (out_fmt == 0) orc_output_float(p, yuv, yuv + ca, yuv + 2 * ca, (float *)out);
    else if (out_fmt == 1) orc_output_u16(p, yuv, yuv + ca, yuv + 2 * ca, (uint16_t *)out);
    else if (out_fmt == 2) orc_output_rgb48(p, yuv, yuv + ca, yuv + 2 * ca, (uint16_t *)out);
    else orc_output_rgbs(p, yuv, yuv + ca, yuv + 2 * ca, (float *)out);
}

/* ------------------------------------------------------------------------ */
/* End-to-end: one frame from a field window, as chd_decode_frame composes    */
/* SourceField::loadFields -> PalColour::decodeFrames -> OutputWriter         */
/* (c_api_decoder.cpp:183-332, 461-479, 497-515).                            */
/* kind 7 = TRANSFORM_2D (window = the frame's 2 fields, start 0, end 2);     */
/* kind 8 = TRANSFORM_3D (window of 14 fields, start 4, end 6).              */
/* out_fmt as write_output.  Optional chroma_out receives                     */
/* the two chroma fields (2 x field area doubles) and yuv_out the component  */
/* frame (3 x field_width x (2*field_height-1) doubles).                      */
/* ------------------------------------------------------------------------ */

static int decode_window(const orc_params *p, int32_t kind, const double *thr_sq, const float *masks,
                         const uint16_t *const *fields, const int32_t *offsets, int32_t nfields, int32_t start,
                         int32_t end, int32_t out_fmt, void *out, double *chroma_out, double *yuv_out)
{
    if (end - start != 2) return -1;
    const size_t fa = (size_t)p->field_width * p->field_height;
    const size_t ca = (size_t)p->field_width * (2 * p->field_height - 1);
    double *chroma = malloc(2 * fa * sizeof(double));
    double *outs[2] = {chroma, chroma + fa};
    if (kind == 6) {
        /* CHD_DEC_PAL_2D: no transform; the decoder filters the composite */
        double *yuv = malloc(3 * ca * sizeof(double));
        orc_pal2d_decode_frame(p, fields[start], offsets[start], fields[start + 1], offsets[start + 1], yuv,
                               yuv + ca, yuv + 2 * ca);
        write_output(p, yuv, ca, out_fmt, out);
        if (chroma_out) memset(chroma_out, 0, 2 * fa * sizeof(double));
        if (yuv_out) memcpy(yuv_out, yuv, 3 * ca * sizeof(double));
        free(yuv);
        free(chroma);
        return 0;
    }
    if (kind == 8) {
        if (start < 4 || nfields - end < 4) { free(chroma); return -2; }
        tpal3d_filter(p, fields, offsets, start, end, thr_sq, masks, outs);
    } else if (kind == 7) {
        orc_tpal2d_filter_fields(p, fields, offsets, nfields, start, end, thr_sq, outs);
    } else {
        free(chroma);
        return -3;
    }
    double *yuv = malloc(3 * ca * sizeof(double));
    orc_palcolour_decode_frame(p, fields[start], offsets[start], fields[start + 1], offsets[start + 1],
                               outs[0], outs[1], yuv, yuv + ca, yuv + 2 * ca);
    write_output(p, yuv, ca, out_fmt, out);
    if (chroma_out) memcpy(chroma_out, chroma, 2 * fa * sizeof(double));
    if (yuv_out) memcpy(yuv_out, yuv, 3 * ca * sizeof(double));
    free(yuv);
    free(chroma);
    return 0;
}

int orc_decode_window(const orc_params *p, int32_t kind, const double *thr_sq, const uint16_t *const *fields,
                      const int32_t *offsets, int32_t nfields, int32_t start, int32_t end, int32_t out_fmt,
                      void *out, double *chroma_out, double *yuv_out)
{
    return decode_window(p, kind, thr_sq, NULL, fields, offsets, nfields, start, end, out_fmt, out, chroma_out,
                         yuv_out);
}

/* The forward half of tpal3d_filter for one Z tile (forwardFFTTile,
 * transform_pal_3d.cpp:168-202) on every tile of its lattice, for fields
 * given as doubles (the composite, or a true chroma signal): tile origins
 * as tpal3d_filter's loops, lines outside the active lines or of the other
 * field parity filled with `fill` (black16bIre for a composite, 0 for
 * chroma; a constant line adds nothing at x bins 2..6, the window's x
 * spectrum being bins 0 and +-1).  fields[0..7] are the tile's fields z.
 * out_re/out_im receive, per tile in tile loop order (y then x), the bins
 * x = 2..6 as [k = x - 2][ky][kz] (5 x 32 x 8).                           */
int orc_tile_spectra3d(const orc_params *p, const double *const *fields, const int32_t *offsets, double fill,
                       double *out_re, double *out_im)
{
    const int fw = p->field_width;
    double (*win)[YT3][XT3] = malloc(sizeof(double) * ZT3 * YT3 * XT3);
    for (int z = 0; z < ZT3; z++)
        for (int y = 0; y < YT3; y++)
            for (int x = 0; x < XT3; x++)
                win[z][y][x] = orc_window(z, ZT3) * orc_window(y, YT3) * orc_window(x, XT3);
    double *real = malloc(sizeof(double) * ZT3 * YT3 * XT3);
    double *sre = malloc(sizeof(double) * ZT3 * YT3 * XT3);
    double *sim = malloc(sizeof(double) * ZT3 * YT3 * XT3);
    double *h_re = malloc(sizeof(double) * ZT3 * YT3 * XC3);
    double *h_im = malloc(sizeof(double) * ZT3 * YT3 * XC3);
    const int first = p->first_active_frame_line, last = p->last_active_frame_line;
    const int avs = p->active_video_start, ave = p->active_video_end;
    size_t t = 0;
    for (int tileY = first - YT3 / 2; tileY <= last; tileY += YT3 / 2) {
        for (int tileX = avs - XT3 / 2; tileX < ave; tileX += XT3 / 2, t++) {
            const int startY = imax(first - tileY, 0);
            const int endY = imin((last - tileY) + 1, YT3);
            for (int z = 0; z < ZT3; z++) {
                for (int y = 0; y < YT3; y++) {
                    double *dst = real + (z * YT3 + y) * XT3;
                    if (y < startY || y >= endY || ((tileY + y) % 2) != offsets[z]) {
                        for (int x = 0; x < XT3; x++) dst[x] = fill * win[z][y][x];
                        continue;
                    }
                    const double *b = fields[z] + (size_t)((tileY + y) / 2) * fw;
                    for (int x = 0; x < XT3; x++) {
                        const int sx = tileX + x;
                        dst[x] = ((sx >= 0 && sx < fw) ? b[sx] : fill) * win[z][y][x];
                    }
                }
            }
            orc_r2c3d(real, ZT3, YT3, XT3, h_re, h_im, sre, sim);
            double *ore = out_re + t * 5 * YT3 * ZT3, *oim = out_im + t * 5 * YT3 * ZT3;
            for (int k = 0; k < 5; k++)
                for (int ky = 0; ky < YT3; ky++)
                    for (int kz = 0; kz < ZT3; kz++) {
                        ore[(k * YT3 + ky) * ZT3 + kz] = h_re[(kz * YT3 + ky) * XC3 + k + 2];
                        oim[(k * YT3 + ky) * ZT3 + kz] = h_im[(kz * YT3 + ky) * XC3 + k + 2];
                    }
        }
    }
    free(win);
    free(real);
    free(sre);
    free(sim);
    free(h_re);
    free(h_im);
    return 0;
}

/* 3D decode of one frame window with a soft mask per tile (masks: 768
 * floats per tile of the window's two Z tiles, in tile loop order). */
int orc_decode_window_masked(const orc_params *p, const float *masks, const uint16_t *const *fields,
                             const int32_t *offsets, int32_t nfields, int32_t start, int32_t end, int32_t out_fmt,
                             void *out, double *chroma_out)
{
    return decode_window(p, 8, NULL, masks, fields, offsets, nfields, start, end, out_fmt, out, chroma_out, NULL);
}

/* Decode frames of a regularly paired sequence (first field first, field 2f
 * and 2f+1 of `seq` form frame f, 0-based), with black padding outside the
 * sequence as SourceField::loadFields does (source_field.cpp:87-158).  Frames
 * run in parallel over `threads` OpenMP threads, one decoder per thread, the
 * way chd_decode_frames_async gives each worker its own decoder
 * (c_api_decoder.cpp:1093-1152).  out receives n consecutive output frames.  */
int orc_decode_sequence(const orc_params *p, int32_t kind, const double *thr_sq, const uint16_t *seq,
                        int32_t n_seq_fields, const int32_t *frames, int32_t n, int32_t out_fmt, void *out,
                        int32_t threads)
{
    const size_t fa = (size_t)p->field_width * p->field_height;
    const size_t ps = (size_t)p->out_width * p->out_height * 3;
    const int behind = (kind == 8) ? 2 : 0, ahead = (kind == 8) ? 4 : 0;
    const int nwin = 2 * (behind + 1 + ahead);
    uint16_t *blackf = malloc(fa * sizeof(uint16_t));
    for (size_t i = 0; i < fa; i++) blackf[i] = (uint16_t)p->black16b;
    int rc_all = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int i = 0; i < n; i++) {
        const uint16_t *win[2 * 7];
        int32_t offs[2 * 7];
        for (int w = 0; w < nwin; w++) {
            const int fi = 2 * (frames[i] - behind) + w;
            win[w] = (fi >= 0 && fi < n_seq_fields) ? seq + (size_t)fi * fa : blackf;
            offs[w] = w % 2;
        }
        void *o = (out_fmt == 0 || out_fmt == 3) ? (void *)((float *)out + ps * i) : (void *)((uint16_t *)out + ps * i);
        const int rc = orc_decode_window(p, kind, thr_sq, win, offs, nwin, 2 * behind, 2 * behind + 2, out_fmt, o,
                                         NULL, NULL);
        if (rc != 0) rc_all = rc;
    }
    free(blackf);
    return rc_all;
}
