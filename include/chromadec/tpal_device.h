/* libchromadec-mi355x: device-resident Transform-PAL decode.
 *
 * The thin extern "C" HIP layer under the public ABI.  It has no
 * counterpart in the reference's public headers: it replaces the reference's
 * internal seam Decoder::decodeFrames (src/decoders/decoder_base.h:69-71),
 * i.e. TransformPal{2D,3D}::filterFields (src/decoders/palcolour/
 * transform_pal.h:65-66) + PalColour::decodeFrames (palcolour.cpp:279-304)
 * + OutputWriter::convertToFloat / convert (src/output/output_writer.cpp:
 * 340-486, 670-727), for callers whose fields already sit in GPU memory
 * (batch transcoders, benchmarks, other GPU stages).  chd_decode_frame is
 * built on it.
 *
 * Every pointer named d_* is device memory of the plan's device; `stream`
 * is a hipStream_t (NULL = the default stream).  A plan is not thread-safe:
 * use one per host thread.  Results are identical to chd_decode_frame on the
 * same fields (the per-frame Z-tile lattice of the ABI is kept). */
#ifndef CHROMADEC_TPAL_DEVICE_H
#define CHROMADEC_TPAL_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#include <chromadec/errors.h>
#include <chromadec/types.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chd_tpal_plan chd_tpal_plan_t;

typedef struct chd_tpal_params {
    int32_t kind;                        /* CHD_DEC_PAL_2D (6), CHD_DEC_TRANSFORM_2D (7) or _3D (8) */
    int32_t field_width, field_height;
    int32_t active_video_start, active_video_end;            /* samples, half-open */
    int32_t first_active_frame_line, last_active_frame_line; /* inclusive */
    int32_t colour_burst_start, colour_burst_end;
    int32_t black_16b_ire, white_16b_ire;
    double  fsc_hz, sample_rate_hz;
    double  threshold;                   /* used when thresholds is NULL (default 0.4) */
    const double *thresholds;            /* 768 (3D) or 80 (2D) per-bin thresholds, or NULL */
    double  chroma_gain, chroma_phase_deg;
    double  ecb_from_u, ecr_from_v;      /* composite U/V -> E'Cb/E'Cr per unit signal range */
    int32_t pixel_format;                /* CHD_PIXEL_YUV444PS, _YUV444P16, _GRAYS, _GRAY16, _RGB48 or _RGBS */
    int32_t clamp;                       /* chd_clamp_t */
    int32_t out_width, out_height;       /* output frame incl. padding */
    int32_t pad_left, pad_top;           /* position of the active picture */
    int32_t max_batch_frames;            /* frames per internal launch; 0 = 64 */
    int32_t device;                      /* HIP device ordinal */
    double  kr, kb;                      /* luma coefficients of the colour-difference
                                            precision (R'G'B' output); 0 = 0.299, 0.114 */
    double  luma_nr_level;               /* luma coring in IRE (PalColour yNRLevel,
                                            palcolour.cpp:411-448); 0 = off */
} chd_tpal_params_t;

/* Build a plan: constant tables, thresholds and a workspace sized for
 * max_batch_frames.  CHD_E_UNSUPPORTED without a usable HIP device,
 * CHD_E_INVALID_ARG for geometry the kernels do not cover. */
chd_status_t chd_tpal_plan_create(const chd_tpal_params_t *params, chd_tpal_plan_t **out);
void         chd_tpal_plan_free(chd_tpal_plan_t *plan);

/* uint16 elements per packed field, and bytes per output frame. */
size_t       chd_tpal_packed_field_elems(const chd_tpal_plan_t *plan);
size_t       chd_tpal_frame_bytes(const chd_tpal_plan_t *plan);

/* Pack n_fields fields stored as in a .tbc file (field_height rows of
 * field_width samples, back to back) into the plan's aligned field layout. */
chd_status_t chd_tpal_pack_fields(chd_tpal_plan_t *plan, const uint16_t *d_src, int64_t n_fields,
                                  uint16_t *d_packed, void *stream);

/* A run of frames of one capture held as packed fields: frame k is fields
 * 2k (first field) and 2k+1.  Frames outside [0, n_frames) read as black,
 * as the reference pads the look-around window (source_field.cpp:155-158). */
typedef struct chd_tpal_sequence {
    const uint16_t *d_fields;
    int64_t n_frames;
    int64_t first_frame;
    int64_t count;
} chd_tpal_sequence_t;

/* Decode every listed run of every sequence into d_out: the frames of
 * seqs[0] first, then seqs[1], ..., each chd_tpal_frame_bytes() long.
 * Each frame equals chd_decode_frame of the same frame index.  Enqueued on
 * `stream`; returns without waiting. */
chd_status_t chd_tpal_decode(chd_tpal_plan_t *plan, const chd_tpal_sequence_t *seqs, int32_t n_seqs,
                             void *d_out, void *stream);

/* chd_tpal_decode without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory at decode
 * time (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The overlap-add partials, job lists and (with luma
 * coring) uncored planes live in one static device array of
 * CHD_TPAL_STATIC_SCRATCH_MIB MiB per GPU, held until the process ends, so
 * frames are decoded in chunks of as many frames as fit (4 for 3D PAL,
 * 12 for 2D).  Output equals chd_tpal_decode's bit for bit.  The plan still
 * comes from chd_tpal_plan_create (which allocates) and must have no learned
 * mask (CHD_E_UNSUPPORTED); a geometry whose single frame needs more than the
 * array is rejected with CHD_E_INVALID_ARG.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (any plan, any stream)
 * race, so serialise them. */
#define CHD_TPAL_STATIC_SCRATCH_MIB 64
chd_status_t chd_tpal_decode_static(chd_tpal_plan_t *plan, const chd_tpal_sequence_t *seqs, int32_t n_seqs,
                                    void *d_out, void *stream);

/* Demodulate frames whose chroma was separated elsewhere: the second half
 * of the reference's decode, PalColour::decodeFrames with prefiltered
 * chroma (palcolour.cpp:279-335) and OutputWriter, without the transform.
 * d_chroma holds for fields 2k, 2k+1 of every frame field_height rows of
 * field_width floats: chroma in 16-bit sample units, as
 * TransformPal::filterFields outputs it (transform_pal.h:60-66); only the
 * active samples are read.  With the true chroma of a synthetic capture this
 * gives the ideal-separation picture a separation is measured against.
 * Frames are fields 2k, 2k+1 of d_fields (packed); output as chd_tpal_decode. */
chd_status_t chd_tpal_decode_separated(chd_tpal_plan_t *plan, const uint16_t *d_fields, const float *d_chroma,
                                       int64_t n_frames, void *d_out, void *stream);

/* A fixed batch: every listed frame run of every sequence decoded by one
 * tile-transform launch (persistent workgroups over all slabs) and one
 * demodulation launch, captured once in a HIP graph and replayed by
 * chd_tpal_batch_run.  Job lists and the overlap-add workspace for the whole
 * batch live in device memory; the sequences' fields and d_out must stay at
 * the same addresses for the life of the batch (their contents may change).
 * With autotune != 0 (3D plans) the strip length (tiles per work item: 4, 8,
 * 12, 16, 24) and the grid (one workgroup per item, or 2-12 persistent
 * workgroups per CU) are chosen by timing each on this batch; without it the
 * strip length comes from the launch size and the grid is one workgroup per
 * item.  A plan with a learned
 * mask network (chd_tpal_set_mask_net) is captured with it.  The plan must
 * outlive the batch; use one batch per stream at a time. */
typedef struct chd_tpal_batch chd_tpal_batch_t;
chd_status_t chd_tpal_batch_create(chd_tpal_plan_t *plan, const chd_tpal_sequence_t *seqs, int32_t n_seqs,
                                   void *d_out, int32_t autotune, chd_tpal_batch_t **out);
chd_status_t chd_tpal_batch_run(chd_tpal_batch_t *batch, void *stream);
/* Chosen launch shape: tiles per work item, workgroups of the tile launch
 * (0 = one per item), and the tile-launch time it measured (ms, 0 without
 * autotuning). */
chd_status_t chd_tpal_batch_config(const chd_tpal_batch_t *batch, int32_t *strip, int32_t *groups,
                                   double *tile_ms);
/* Re-capture with a given shape (e.g. one an earlier autotune chose). */
chd_status_t chd_tpal_batch_set_shape(chd_tpal_batch_t *batch, int32_t strip, int32_t groups);
void         chd_tpal_batch_free(chd_tpal_batch_t *batch);

/* Kernel-level timing: milliseconds in the tile-transform and demodulation
 * kernels, from HIP events recorded on the decode stream without waiting.
 * enable: 0 off; 1 the totals cover the last decode call on this plan;
 * 2 the totals run over every decode call since chd_tpal_set_timing.
 * chd_tpal_last_timing waits for the recorded kernels, then reports. */
chd_status_t chd_tpal_set_timing(chd_tpal_plan_t *plan, int enable);
chd_status_t chd_tpal_last_timing(chd_tpal_plan_t *plan, double *tile_ms, double *demod_ms,
                                  int32_t *n_tile_launches, int32_t *n_demod_launches);

#ifdef __cplusplus
}
#endif
#endif
