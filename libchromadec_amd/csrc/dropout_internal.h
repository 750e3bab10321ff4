// Dropout concealment on host field data, before the chroma decode.
//
// Follows DropoutCorrector (src/dropout/dropout_corrector.cpp:41-530) for a
// single source: each flagged run of samples is replaced from the nearest
// line without an overlapping dropout that has the same subcarrier phase
// (4 field lines away in PAL, or the other field at offset -3 / -1), with
// low frequencies from the nearest line of any phase and high frequencies
// from the nearest same-phase line when the two differ.
#pragma once

#include <stdint.h>

#include <vector>

namespace chd {
namespace dropout {

struct Params {
    int32_t field_width = 0, field_height = 0;
    int32_t colour_burst_end = 0;
    int32_t active_video_end = 0;          // post-padding, half-open
    int32_t first_active_frame_line = 0;   // post-padding, inclusive
    int32_t last_active_frame_line = 0;
    int32_t system = 0;                    // 0 PAL, 1 NTSC, 2 PAL-M, 3 SECAM (chd::System)
};

struct Run {
    int32_t startx = 0, endx = 0;   // half-open
    int32_t field_line = 0;         // 1-based
};

struct Stats {
    int32_t corrected = 0;
    int32_t failed = 0;
    int64_t total_distance = 0;
};

// Correct one frame in place.  first/second are the frame's first and
// second field (field_height rows of field_width samples); d1/d2 their
// dropout runs as recorded in the metadata.
void correct_frame(const Params &p, uint16_t *first, uint16_t *second, const std::vector<Run> &d1,
                   const std::vector<Run> &d2, bool overcorrect, bool intra_field, Stats *stats);

}  // namespace dropout
}  // namespace chd
