/* libchromadec-mi355x public ABI: version and features
 * (reference include/chromadec/version.h.in).  The version tracks the
 * reference release whose ABI this library implements. */
#ifndef CHROMADEC_VERSION_H
#define CHROMADEC_VERSION_H

#ifdef __cplusplus
extern "C" {
#endif

#define CHROMADEC_VERSION_MAJOR 0
#define CHROMADEC_VERSION_MINOR 2
#define CHROMADEC_VERSION_PATCH 1
#define CHROMADEC_VERSION_STRING "0.2.1-mi355x"

const char *chd_version_string(void);
void        chd_version(int *major, int *minor, int *patch);

/* "rocm" (always 1: decoding runs on HIP); every other
 * feature name of the reference ("nn", "onnxruntime", "coreml", "cuda",
 * "fftw", "hvd", "sqlite") answers 0. */
int         chd_has_feature(const char *feature);

#ifdef __cplusplus
}
#endif
#endif
