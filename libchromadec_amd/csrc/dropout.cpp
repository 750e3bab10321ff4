// Dropout concealment (DropoutCorrector, src/dropout/dropout_corrector.cpp,
// single source).  Runs on the host on a frame's two fields before they are
// uploaded: the work is a few line copies per frame, and the result must not
// depend on scheduling (later corrections read lines that earlier ones wrote,
// as in the reference).
#include "dropout_internal.h"

#include <algorithm>
#include <cstdlib>

namespace chd {
namespace dropout {

namespace {

enum Location { kVisible, kBurst, kUnknown };

struct Loc {
    int32_t field_line, startx, endx;
    Location where;
};

struct Replacement {
    bool same_field = true;
    int32_t field_line = -1;
    double quality = -1.0;
    int32_t distance = 0;
};

// 2.2 MHz 5-tap luma low-pass at 4fSC for PAL and SECAM, 1.8 MHz for NTSC
// and PAL-M (luma_filters.cpp:36-52); samples outside the line count as 0
// and the sum is truncated to uint16 (firfilter.h apply).
constexpr double kPalLuma[5] = {0.03283437, 0.23959832, 0.45513461, 0.23959832, 0.03283437};
constexpr double kNtscLuma[5] = {0.03275786, 0.23955702, 0.45537024, 0.23955702, 0.03275786};

void luma_lowpass(const Params &p, const uint16_t *in, uint16_t *out)
{
    const double *c = (p.system == 0 || p.system == 3) ? kPalLuma : kNtscLuma;
    const int n = p.field_width;
    for (int i = 0; i < n; ++i) {
        double v = 0.0;
        for (int j = 0; j < 5; ++j) {
            const int k = i - 2 + j;
            if (k >= 0 && k < n) v += c[j] * in[k];
        }
        out[i] = (uint16_t)v;
    }
}

std::vector<Loc> populate(const Params &p, const std::vector<Run> &runs, bool overcorrect)
{
    std::vector<Loc> out;
    for (const Run &r : runs) {
        Loc l{r.field_line, r.startx, r.endx, kUnknown};
        if (l.field_line < 1 || l.field_line > p.field_height) continue;
        if (overcorrect) {
            // widen by 24 samples each side (dropout_corrector.cpp:190-201)
            l.startx = l.startx > 24 ? l.startx - 24 : 0;
            l.endx = l.endx < p.field_width - 24 ? l.endx + 24 : p.field_width;
        }
        out.push_back(l);
    }
    return out;
}

// Classify runs as colour burst or visible line, splitting runs that cross
// the end of the burst and clipping runs at the end of the active line
// (dropout_corrector.cpp:208-268).
void set_locations(const Params &p, std::vector<Loc> &d)
{
    int splits;
    do {
        splits = 0;
        const size_t n = d.size();
        for (size_t i = 0; i < n; ++i) {
            if (d[i].startx <= p.colour_burst_end) {
                d[i].where = kBurst;
                if (d[i].endx > p.colour_burst_end) {
                    d.push_back(Loc{d[i].field_line, p.colour_burst_end + 1, d[i].endx, kBurst});
                    d[i].endx = p.colour_burst_end;
                    ++splits;
                }
            } else if (d[i].startx <= p.active_video_end) {
                d[i].where = kVisible;
                if (d[i].endx > p.active_video_end) {
                    d[i].endx = p.active_video_end;
                    ++splits;
                }
            }
        }
    } while (splits != 0);
}

// Scan a field from target line + offset in steps for the first line without
// an overlapping dropout (dropout_corrector.cpp:409-455).
void find_potential(const Params &p, const Loc &target, const std::vector<Loc> &src, bool same_field, int32_t offset,
                    int32_t step, std::vector<Replacement> &cands)
{
    const int32_t first = (p.first_active_frame_line + 1) / 2;
    const int32_t last = p.last_active_frame_line / 2;
    int32_t line = target.field_line + offset;
    if (line - 1 < first || line - 1 > last) return;
    while (line - 1 >= first && line <= last) {
        bool overlap = false;
        for (const Loc &s : src) {
            if (s.field_line == line && target.endx - s.startx >= 0 && s.endx - target.startx >= 0) {
                overlap = true;
                break;
            }
        }
        if (!overlap) {
            Replacement r;
            r.same_field = same_field;
            r.field_line = line;
            r.quality = 0.0;   // a single source has no quality ordering
            cands.push_back(r);
            return;
        }
        line += step;
    }
}

// Nearest usable line in frame-line distance (dropout_corrector.cpp:273-402).
Replacement find_replacement(const Params &p, const std::vector<Loc> &mine, const std::vector<Loc> &other,
                             size_t idx, bool this_first, bool match_phase, bool burst, bool intra)
{
    int32_t step, other_offset;
    if (!match_phase) {
        step = 1;
        other_offset = -1;
    } else if (p.system == 0 || p.system == 2) {
        // PAL: same subcarrier phase, burst phase and V-switch 4 field lines
        // away; across fields first line N pairs with N-3, second N with N-1
        step = 4;
        other_offset = this_first ? -3 : -1;
    } else if (p.system == 3) {
        step = 2;
        other_offset = -1;
    } else {
        step = 2;
        other_offset = -1;
    }
    std::vector<Replacement> cands;
    const Loc &t = mine[idx];
    find_potential(p, t, mine, true, 0, -step, cands);
    find_potential(p, t, mine, true, step, step, cands);
    if (!burst && !intra) {
        find_potential(p, t, other, false, other_offset, -step, cands);
        find_potential(p, t, other, false, other_offset + step, step, cands);
    }
    Replacement best;
    if (!cands.empty()) {
        best.distance = 1000000;
        best.quality = -1.0;
        const int32_t dl = 2 * t.field_line + (this_first ? 0 : 1);
        for (const Replacement &c : cands) {
            const int32_t sl = 2 * c.field_line + (c.same_field ? (this_first ? 0 : 1) : (this_first ? 1 : 0));
            const int32_t dist = std::abs(dl - sl);
            if (dist < best.distance || (dist == best.distance && c.quality > best.quality)) {
                best = c;
                best.distance = dist;
            }
        }
    }
    return best;
}

// Copy or blend the replacement into the run (dropout_corrector.cpp:458-530).
void conceal(const Params &p, const Loc &d, const Replacement &r, const Replacement &cr, uint16_t *mine,
             const uint16_t *other)
{
    if (r.field_line == -1) return;
    const int32_t w = p.field_width;
    const uint16_t *src = (r.same_field ? mine : other) + (size_t)(r.field_line - 1) * w;
    uint16_t *dst = mine + (size_t)(d.field_line - 1) * w;
    if (cr.field_line == -1 || (d.field_line == r.field_line && d.field_line == cr.field_line)) {
        for (int32_t x = d.startx; x < d.endx; ++x) dst[x] = src[x];
        return;
    }
    std::vector<uint16_t> lf(w);
    luma_lowpass(p, src, lf.data());
    for (int32_t x = d.startx; x < d.endx; ++x) dst[x] = lf[x];
    const uint16_t *chroma = (cr.same_field ? mine : other) + (size_t)(cr.field_line - 1) * w;
    luma_lowpass(p, chroma, lf.data());
    for (int32_t x = d.startx; x < d.endx; ++x) dst[x] = (uint16_t)(dst[x] + (int)chroma[x] - (int)lf[x]);
}

void correct_field(const Params &p, const std::vector<Loc> &mine, const std::vector<Loc> &other, uint16_t *data,
                   const uint16_t *other_data, bool this_first, bool intra, Stats *stats)
{
    for (size_t i = 0; i < mine.size(); ++i) {
        Replacement r, cr;
        if (mine[i].where == kBurst) {
            r = find_replacement(p, mine, other, i, this_first, true, true, intra);
        } else if (mine[i].where == kVisible) {
            r = find_replacement(p, mine, other, i, this_first, false, false, intra);
            cr = find_replacement(p, mine, other, i, this_first, true, false, intra);
        }
        if (stats) {
            if (r.field_line == -1) {
                stats->failed++;
            } else {
                stats->corrected++;
                stats->total_distance += r.distance;
            }
        }
        conceal(p, mine[i], r, cr, data, other_data);
    }
}

}  // namespace

void correct_frame(const Params &p, uint16_t *first, uint16_t *second, const std::vector<Run> &d1,
                   const std::vector<Run> &d2, bool overcorrect, bool intra_field, Stats *stats)
{
    if (d1.empty() && d2.empty()) return;
    std::vector<Loc> l1 = populate(p, d1, overcorrect), l2 = populate(p, d2, overcorrect);
    set_locations(p, l1);
    set_locations(p, l2);
    correct_field(p, l1, l2, first, second, true, intra_field, stats);
    correct_field(p, l2, l1, second, first, false, intra_field, stats);
}

}  // namespace dropout
}  // namespace chd
