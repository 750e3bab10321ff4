/* libchromadec-mi355x: dropout concealment ahead of the chroma decode.
 *
 * Same types and entry points as the reference's include/chromadec/dropout.h
 * (chd_dropout_opts_t, chd_dropout_stats_t, chd_decoder_set_dropout,
 * chd_decoder_get_last_dropout_stats; dropout.h:13-34).  Correction uses the
 * capture's own dropout metadata (the sidecar's dropOuts lists) and a single
 * source: extra sources, dropout-span queries and dropout masks are not part
 * of this build. */
#ifndef CHROMADEC_DROPOUT_H
#define CHROMADEC_DROPOUT_H

#include <stdint.h>

#include <chromadec/errors.h>
#include <chromadec/types.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chd_dropout_opts {
    int enabled;
    int overcorrect;        /* extend dropout boundaries by 24 samples each side */
    int intra_field_only;   /* skip replacement candidates from the other field */
    void *reserved[4];      /* zero-initialise */
} chd_dropout_opts_t;

typedef struct chd_dropout_stats {
    int32_t corrected;      /* runs replaced */
    int32_t failed;         /* runs with no usable replacement line */
    int64_t total_distance; /* sum of frame-line distances of the replacements */
} chd_dropout_stats_t;

/* Dropout options for later decodes on this decoder (may change between
 * decode calls).  Every frame of a decode's look-around window is corrected,
 * so a 3D decode never sees a neighbour's dropouts. */
chd_status_t chd_decoder_set_dropout(chd_decoder_t *d, const chd_dropout_opts_t *opts);

/* Counts for the frame of the most recent chd_decode_frame on this decoder. */
chd_status_t chd_decoder_get_last_dropout_stats(const chd_decoder_t *d, chd_dropout_stats_t *out);

/* No counterpart in the reference's public ABI: conceal the dropouts of frame
 * frame_index (0-based) of `video`, as its metadata lists them, in host
 * copies of the frame's first and second field (field_height rows of
 * field_width samples each), exactly as chd_decode_frame does for a decoder
 * with the video's own active area.  For callers that feed fields to the
 * device decode (chd_tpal_decode) themselves. */
chd_status_t chd_video_correct_dropouts(const chd_video_t *video, int64_t frame_index, const chd_dropout_opts_t *opts,
                                        uint16_t *first_field, uint16_t *second_field, chd_dropout_stats_t *stats);

#ifdef __cplusplus
}
#endif
#endif
