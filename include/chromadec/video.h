/* libchromadec-mi355x public ABI: library lifecycle and composite sources.
 *
 * Replaces the Transform-PAL-relevant part of the reference's
 * include/chromadec/video.h:12-13,31-34,44-50,77-80 (chd_init/shutdown,
 * chd_video_open_composite, chd_video_free, chd_video_get_info and the
 * frame-line/signal-line conversions).  Y/C pairs, CVBS files and extra
 * dropout sources are not part of this build. */
#ifndef CHROMADEC_VIDEO_H
#define CHROMADEC_VIDEO_H

#include <chromadec/errors.h>
#include <chromadec/types.h>

#ifdef __cplusplus
extern "C" {
#endif

chd_status_t chd_init(void);
void         chd_shutdown(void);

/* Open an ld-decode `.tbc` composite capture.  metadata_path_or_null: NULL
 * locates `<path>.json` next to the data; otherwise the path of a `.json`
 * sidecar.  override_or_null may re-declare is_subcarrier_locked and
 * is_second_field_first over the sidecar; standard must be 0 or agree with
 * the sidecar's.  CHD_E_FILE_NOT_FOUND if the data or sidecar is missing,
 * CHD_E_METADATA_CORRUPT if the sidecar does not parse, CHD_E_FORMAT_UNSUPPORTED
 * for a SQLite (`.db`) sidecar. */
chd_status_t chd_video_open_composite(const char *path,
                                      const char *metadata_path_or_null,
                                      const chd_video_params_t *override_or_null,
                                      chd_video_t **out);

void         chd_video_free(chd_video_t *v);

chd_status_t chd_video_get_info(const chd_video_t *v, chd_video_info_t *out);

/* Woven frame line (0..2*field_height-2) <-> field-sequential signal line
 * (1..2*field_height-1); CHD_E_OUT_OF_RANGE outside those ranges. */
chd_status_t chd_video_frame_line_to_signal_line(const chd_video_t *v,
                                                  int32_t frame_line, int32_t *out);
chd_status_t chd_video_signal_line_to_frame_line(const chd_video_t *v,
                                                  int32_t signal_line, int32_t *out);

#ifdef __cplusplus
}
#endif
#endif
