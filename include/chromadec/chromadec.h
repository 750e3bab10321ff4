/* libchromadec-mi355x public umbrella header. */
#ifndef CHROMADEC_CHROMADEC_H
#define CHROMADEC_CHROMADEC_H

#include <chromadec/version.h>
#include <chromadec/errors.h>
#include <chromadec/log.h>
#include <chromadec/types.h>
#include <chromadec/video.h>
#include <chromadec/decoder.h>
#include <chromadec/frame.h>
#include <chromadec/dropout.h>
#include <chromadec/pipeline.h>

#endif
