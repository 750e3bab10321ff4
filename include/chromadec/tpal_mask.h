/* libchromadec-mi355x: learned carrier-mirror mask for CHD_DEC_TRANSFORM_3D.
 *
 * No counterpart in the reference.  The network takes the place of the
 * hard keep/discard rule of TransformPal3D::applyFilter
 * (src/decoders/palcolour/transform_pal_3d.cpp:249-333): for each of the 768
 * visits of that rule (v = (z*32 + y)*3 + (x-2), the thresholds-file order)
 * it predicts a soft mask in [0, 1] from the magnitudes of the visit's bin
 * and its mirror about the U/V carrier and their neighbours.  The plain
 * threshold stays the fallback (no network set, or no weights) and the
 * yardstick the learned form is measured against.
 *
 * Network: three 3x3x3 convolutions over the visit grid (8 x 32 x 3,
 * circular in z and y, zero-padded in x), 4 -> 32 -> 32 -> 1 channels,
 * ReLU, ReLU, sigmoid; BF16 on the matrix cores with FP32 accumulation.
 * Parameters are CHD_MASK_NET_PARAMS floats:
 *   W1[27][4][32], b1[32], W2[27][32][32], b2[32], W3[27][32], b3
 * with tap t = (dz+1)*9 + (dy+1)*3 + (dx+1).
 *
 * Spectra are the 3D tile kernel's bins x = 2..6: per tile
 * CHD_TPAL_TILE_BINS complex floats [k = x-2][ky][kz] (chd_tpal_slab_spectra).
 * Every pointer named d_* is device memory of the network's device. */
#ifndef CHROMADEC_TPAL_MASK_H
#define CHROMADEC_TPAL_MASK_H

#include <stdint.h>

#include <chromadec/errors.h>
#include <chromadec/tpal_device.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHD_MASK_NET_PARAMS 32033
#define CHD_TPAL_TILE_BINS 1280
#define CHD_TPAL_TILE_VISITS 768

typedef struct chd_mask_net chd_mask_net_t;

/* Network state on `device` with training working memory for max_tiles
 * tiles per gradient call (about 0.2 MB per tile; 1 for a network used only
 * to decode).  Parameters start at 0 (mask 0.5 everywhere) until set or
 * loaded. */
chd_status_t chd_mask_net_create(int32_t device, int64_t max_tiles, chd_mask_net_t **out);
void         chd_mask_net_free(chd_mask_net_t *net);

/* Host copies of the parameters; n must be CHD_MASK_NET_PARAMS. */
chd_status_t chd_mask_net_set_params(chd_mask_net_t *net, const float *params, int64_t n);
chd_status_t chd_mask_net_get_params(chd_mask_net_t *net, float *params, int64_t n);

/* Weights file: the 8 bytes "CHDMASK1", a little-endian int32 parameter
 * count, then the parameters as little-endian float32. */
chd_status_t chd_mask_net_load(chd_mask_net_t *net, const char *path);
chd_status_t chd_mask_net_save(chd_mask_net_t *net, const char *path);

/* Soft mask of n_tiles tiles: d_mask[tile][768].  Any n_tiles (inference
 * needs no working memory: one fused kernel, activations on chip).
 * Enqueued on `stream`. */
chd_status_t chd_mask_net_forward(chd_mask_net_t *net, const void *d_spectra, int64_t n_tiles, float *d_mask,
                                  void *stream);

/* Training: gradient of loss_scale * sum over tiles and visits of
 * (w m^2 - 2 num m), where for the bins a visit owns (its own, plus its
 * partner for x = 2, 3) w = sum |F|^2 and num = sum Re(conj(C) F), F the
 * composite's spectrum and C the spectrum of the true chroma of the same
 * tile; this is the squared error sum |m F - C|^2 up to a constant.
 * d_grad receives CHD_MASK_NET_PARAMS floats; *loss (if not NULL) the loss,
 * which waits for the stream.  n_tiles <= max_tiles. */
chd_status_t chd_mask_net_gradient(chd_mask_net_t *net, const void *d_spectra, const void *d_truth_spectra,
                                   int64_t n_tiles, float loss_scale, float *d_grad, double *loss, void *stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the parameters. */
chd_status_t chd_mask_net_adam_step(chd_mask_net_t *net, const float *d_grad, float lr, void *stream);

/* Spectra of every 3D tile of n_slabs slabs of packed fields: slab j is the
 * 8 fields starting at field 2 * (first_slab + j) (fields outside
 * [0, n_fields) read as black), tiles in (tile row, tile column) order,
 * chd_tpal_tiles_per_slab() per slab, CHD_TPAL_TILE_BINS complex floats each.
 * Fields 2k and 2k+1 are the first and second field of frame k. */
int64_t      chd_tpal_tiles_per_slab(const chd_tpal_plan_t *plan);
chd_status_t chd_tpal_slab_spectra(chd_tpal_plan_t *plan, const uint16_t *d_fields, int64_t n_fields,
                                   int64_t first_slab, int32_t n_slabs, void *d_spectra, void *stream);

/* Decode CHD_DEC_TRANSFORM_3D with the learned mask in place of the
 * threshold (net = NULL restores the threshold).  The plan keeps a pointer;
 * the network must outlive its use.  A network's working memory serves one
 * call at a time: plans decoding concurrently need a network each. */
chd_status_t chd_tpal_set_mask_net(chd_tpal_plan_t *plan, chd_mask_net_t *net);

#ifdef __cplusplus
}
#endif
#endif
