/* libchromadec-mi355x public ABI: decoders.
 *
 * Same kinds, option names, signatures and semantics as the reference's
 * include/chromadec/decoder.h:13-104 for the PAL kinds.  CHD_DEC_PAL_2D (6),
 * CHD_DEC_TRANSFORM_2D (7) and CHD_DEC_TRANSFORM_3D (8) decode on the GPU, and
 * CHD_DEC_AUTO resolves to CHD_DEC_PAL_2D for a PAL source; CHD_DEC_NONE
 * commits for geometry and dropout queries only.  Every other kind is
 * rejected at commit with CHD_E_DECODER_UNKNOWN. */
#ifndef CHROMADEC_DECODER_H
#define CHROMADEC_DECODER_H

#include <chromadec/enum.h>
#include <chromadec/errors.h>
#include <chromadec/types.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef CHD_ENUM(chd_decoder_kind) {
    CHD_DEC_AUTO              = 0,
    CHD_DEC_MONO              = 1,
    CHD_DEC_NTSC_1D           = 2,
    CHD_DEC_NTSC_2D           = 3,
    CHD_DEC_NTSC_3D           = 4,
    CHD_DEC_NTSC_3D_NO_ADAPT  = 5,
    CHD_DEC_PAL_2D            = 6,
    CHD_DEC_TRANSFORM_2D      = 7,
    CHD_DEC_TRANSFORM_3D      = 8,
    CHD_DEC_NN_TRANSFORM3D    = 9,
    CHD_DEC_LDZEUG_COLOR_CNN  = 10,
    CHD_DEC_LDZEUG_LUMA_SEP   = 11,
    CHD_DEC_LDZEUG_LUMA_SEP_FRAME = 12,
    CHD_DEC_NONE                  = 13,
    CHD_DEC_SECAM                 = 20,
    CHD_DEC_HVD_2D                = 30,
    CHD_DEC_HVD_3D                = 31
} chd_decoder_kind_t;

chd_status_t chd_decoder_create(chd_video_t *v, chd_decoder_kind_t kind, chd_decoder_t **out);
void         chd_decoder_free(chd_decoder_t *d);

/* Typed option setters: CHD_E_INVALID_ARG if the option is not meaningful for
 * the kind or the decoder is already committed.  Must not overlap
 * chd_decode_frame on the same decoder. */
chd_status_t chd_decoder_set_option_f64(chd_decoder_t *d, const char *name, double v);
chd_status_t chd_decoder_set_option_i32(chd_decoder_t *d, const char *name, int32_t v);
chd_status_t chd_decoder_set_option_bool(chd_decoder_t *d, const char *name, int v);
chd_status_t chd_decoder_set_option_str(chd_decoder_t *d, const char *name, const char *v);
chd_status_t chd_decoder_has_option(const chd_decoder_t *d, const char *name);

/* Validate the options, resolve the output framing and build the GPU decode
 * plan.  Required before chd_decode_frame; a second call is a no-op.
 * CHD_E_UNSUPPORTED when no HIP device is usable. */
chd_status_t chd_decoder_commit(chd_decoder_t *d);

chd_status_t chd_decoder_get_output_info(const chd_decoder_t *d, chd_output_info_t *out);

/* Option names (decoder.h:58-88 of the reference). */
#define CHD_OPT_CHROMA_GAIN                 "chroma_gain"               /* f64 */
#define CHD_OPT_CHROMA_PHASE_DEG            "chroma_phase_deg"          /* f64 */
#define CHD_OPT_LUMA_NR_LEVEL               "luma_nr_level"             /* f64, IRE; 0 = off (luma coring) */
#define CHD_OPT_PADDING_MULTIPLE            "padding_multiple"          /* i32, 1..1024 */
#define CHD_OPT_REVERSE_FIELD_ORDER         "reverse_field_order"       /* bool */
#define CHD_OPT_TRANSFORM_THRESHOLD         "transform_threshold"       /* f64, default 0.4 */
#define CHD_OPT_TRANSFORM_THRESHOLDS_FILE   "transform_thresholds_file" /* str */
#define CHD_OPT_FIRST_ACTIVE_SAMPLE         "first_active_sample"       /* i32, inclusive */
#define CHD_OPT_LAST_ACTIVE_SAMPLE          "last_active_sample"        /* i32, inclusive */
#define CHD_OPT_FIRST_ACTIVE_FRAME_LINE     "first_active_frame_line"   /* i32, inclusive */
#define CHD_OPT_LAST_ACTIVE_FRAME_LINE      "last_active_frame_line"    /* i32, inclusive */
#define CHD_OPT_OUTPUT_FORMAT               "output_format"             /* str: "yuv444p16"|"yuv444ps"|"gray16"|"grays"|"rgb48"|"rgbs" */
#define CHD_OPT_OUTPUT_CLAMP                "output_clamp"              /* str: "none"|"legal_rgb_sdr"|"legal_rgb_hdr"|"legal_ycbcr_bt601" */
#define CHD_OPT_COLOR_DIFFERENCE_PRECISION  "color_difference_precision"  /* str: "classic"|"modern" */
#define CHD_OPT_BROADCAST_SCALING_PRECISION "broadcast_scaling_precision" /* str: "classic"|"modern"|"scientific" */
#define CHD_OPT_OUTPUT_Y4M_HEADERS          "output_y4m_headers"        /* bool */
#define CHD_OPT_THREAD_COUNT                "thread_count"              /* i32, 0 = auto */
/* Build-specific: spectral mask for CHD_DEC_TRANSFORM_3D.  "threshold" (the
 * reference's hard carrier-mirror threshold, default) or "learned" (the
 * trained soft mask, with the threshold kept as its fallback). */
#define CHD_OPT_TRANSFORM_MASK              "transform_mask"            /* str */
/* Build-specific: path of the learned-mask weight file for "learned". */
#define CHD_OPT_TRANSFORM_MASK_WEIGHTS      "transform_mask_weights"    /* str */

/* Decode one frame (0-based index) into a new caller-owned frame. */
chd_status_t chd_decode_frame(chd_decoder_t *d, int64_t frame_index, chd_frame_t **out);

/* Decode `n` frames; `cb` receives each (status, index, frame) on an internal
 * worker thread, in completion order.  Returns after every callback ran. */
typedef void (*chd_frame_done_cb)(void *user, chd_status_t s, int64_t idx, chd_frame_t *f);
chd_status_t chd_decode_frames_async(chd_decoder_t *d,
                                      const int64_t *indices, size_t n,
                                      chd_frame_done_cb cb, void *user,
                                      chd_cancel_t *cancel_or_null);

#ifdef __cplusplus
}
#endif
#endif
