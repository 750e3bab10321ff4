/* libchromadec-mi355x public ABI: enum declaration helper.
 *
 * Same contract as the reference's include/chromadec/enum.h:12-16: every
 * public enum is 4 bytes wide; C++ and C23 see it with a fixed int32_t
 * underlying type, older C sees a plain enum of identical representation. */
#ifndef CHROMADEC_ENUM_H
#define CHROMADEC_ENUM_H

#include <stdint.h>

#if defined(__cplusplus) || (defined(__STDC_VERSION__) && __STDC_VERSION__ >= 202311L)
#define CHD_ENUM(name) enum name : int32_t
#else
#define CHD_ENUM(name) enum name
#endif

#endif
