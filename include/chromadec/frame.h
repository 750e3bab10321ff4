/* libchromadec-mi355x public ABI: decoded frames.
 * Replaces the reference's include/chromadec/frame.h:12-53 for the Y'CbCr
 * and grey pixel formats (the 4:4:0 SECAM queries are not part of this build). */
#ifndef CHROMADEC_FRAME_H
#define CHROMADEC_FRAME_H

#include <chromadec/errors.h>
#include <chromadec/types.h>

#ifdef __cplusplus
extern "C" {
#endif

chd_status_t chd_frame_get_info(const chd_frame_t *f, chd_frame_info_t *out);
chd_status_t chd_frame_get_plane_info(const chd_frame_t *f, chd_plane_t p,
                                      chd_plane_info_t *out);

/* Borrow a 16-bit plane (integer formats).  Owned by the frame, valid until
 * chd_frame_free.  Stride in bytes. */
chd_status_t chd_frame_get_plane(const chd_frame_t *f, chd_plane_t p,
                                  const void **out_data,
                                  ptrdiff_t *out_stride_bytes);

/* Borrow a float plane (CHD_PIXEL_YUV444PS: E'Y 0..1, E'Cb/E'Cr +-0.5;
 * CHD_PIXEL_GRAYS: E'Y only).  Same lifetime rules. */
chd_status_t chd_frame_get_plane_float(const chd_frame_t *f, chd_plane_t p,
                                        const float **out_data,
                                        ptrdiff_t *out_stride_bytes);

void         chd_frame_free(chd_frame_t *f);

#ifdef __cplusplus
}
#endif
#endif
