// Host/device shared layout of the Transform-PAL GPU pipeline.
//
// Device memory layout (all HBM):
//   * Packed field store: a field is field_height rows of `pitch` uint16,
//     sample x of a row at column x + lmargin.  lmargin makes every tile
//     origin 16-byte aligned; margins hold black so reads that run off the
//     stored row see black.
//   * Slab partials (3D): one slab = 8 consecutive fields starting at an
//     even field 2k (frame k).  Its inverse transform feeds output fields
//     z = 0,1 (frame k) and z = 4,5 (frame k+2), so a slab partial holds
//     4 planes x 2 row classes (tile row parity) x field_height rows x
//     `cpitch` floats, column c = x - active_video_start + kColOrigin.
//   * Field partials (2D): 2 row classes per field, same row layout.
//   * Output frames: planes x out_height x out_width (float or uint16).
#pragma once

#include <stdint.h>

#include <hip/hip_runtime_api.h>

namespace tpal {

constexpr int kZT = 8, kYT = 32, kXT = 16;  // 3D tile (fields, frame lines, samples)
constexpr int kYT2 = 16, kXT2 = 32;          // 2D tile (field lines, samples)
constexpr int kStrip3 = 8;                   // owned 3D tiles per workgroup strip
constexpr int kStrip2 = 4;                   // owned 2D tiles per lane-group strip
constexpr int kDemodLines = 8;               // field lines per demod workgroup
constexpr int kMaxFilter = 15;               // max PalColour filterSize supported
// Partial column of sample active_video_start.  3D strips own 64 columns
// from tile column 8 tx - 8: with origin 40 every strip starts on a 128-byte
// line, so no two workgroups (on different XCDs) write one cache line.
constexpr int kColOrigin = 40;

struct Geom {
    int32_t fw, fh;          // field width / height
    int32_t avs, ave, aw;    // active samples [avs, ave), aw = ave - avs
    int32_t first, last;     // active frame lines, inclusive
    int32_t cbs, cbe;        // colour burst [cbs, cbe)
    float black;             // black16bIre
    float thr_uniform;       // squared threshold when all bins share it, else -1
    double thr_uniform_d;    // the same in FP64 (the re-decision's threshold)
    int32_t pitch;           // packed row pitch (uint16 elements)
    int32_t lmargin;         // packed column of sample 0
    int64_t field_elems;     // packed field size (elements)
    int32_t cpitch;          // partial row pitch (floats)
    int64_t part_plane;      // fh * cpitch
    // 3D lattice
    int32_t n_ty3, n_tx3, n_strips3;
    // 2D lattice (per field parity: first line differs)
    int32_t n_tx2, n_strips2;
    // demod
    int32_t filter_size;
    int32_t quad;            // 1: sampleRate == 4 fSC, carrier tables exactly 0/+-1
    int32_t spar;            // column parity on which the sine table is non-zero
    float burst_floor;       // 130000 / 128
    // output
    int32_t out_w, out_h, pad_left, pad_top;
    int32_t pixel_format;    // chd_pixel_format_t
    int32_t clamp_mode;
    float ey_scale, ecb_scale, ecr_scale;
    float chroma_gain_cos, chroma_gain_sin;  // gain * cos/sin(phase)
    // R'G'B' output: E'R = E'Y + rgb[0] V, E'G = E'Y + rgb[1] U + rgb[2] V,
    // E'B = E'Y + rgb[3] U (U, V per unit signal range), and the RGBS clamp
    // box of the clamp mode
    float rgb[4];
    float rgb_lo[3], rgb_hi[3];
    // luma coring level in composite units (PalColour::doYNR nr_y), 0 = off
    float ynr;
};

struct SlabJob {
    const uint16_t *field[kZT];   // nullptr = black field
    int32_t offset[kZT];          // 0 = first field of its frame, 1 = second
    float *partial;               // 4 * 2 * part_plane floats
};

struct FieldJob2D {
    const uint16_t *field;        // nullptr = black
    int32_t offset;
    float *partial;               // 2 * part_plane floats
};

struct FrameJob {
    const uint16_t *field[2];
    int32_t offset[2];
    const float *chroma[2][4];    // per field: partial planes summed in order
    int32_t nchroma;              // 0: CHD_DEC_PAL_2D, the filter runs on the composite itself
    void *out;                    // planes x out_h x out_w
};

// Demodulator filter table: per tap b, kFiltStride floats: the chroma taps
// cfilt[b][0..3] (lines 0, +-2, +-1, +-3), then the luma taps yfilt[b][0..1]
// (lines 0, +-2) used when decoding the raw composite.
constexpr int kFiltStride = 8;

// Constant tables of the tile transforms (device __constant__ copy c_tab).
struct Tables {
    // Window factors: 3D (z, y, x) and 2D (y, x).  The x DFTs use
    // compile-time twiddles.
    float wz[8], wy[32], wx[16];
    float wy2[16], wx2[32];
    // FP64 recheck tables.
    double wz_d[8], wy_d[32], wx_d[16], wy2_d[16], wx2_d[32];
    double cos32_d[32], sin32_d[32];
};


// Launchers (tpal_kernels.hip).
hipError_t upload_tables(const Tables &t);
// 3D tile kernel modes: the hard carrier-mirror mask; tile spectra out
// (bins x = 2..6, kTileBins complex per tile); a soft per-visit mask applied
// to stored spectra.  The paired learned path (half form only): the
// magnitude features of rows ky < 16 out (kTileHalfBins floats per tile),
// then the soft mask applied to the tile transform computed again from the
// composite.
enum { kTilePlain = 0, kTileSpectra = 1, kTileApply = 2, kTileFeatures = 3, kTileApplyFwd = 4 };
constexpr int kTileHalfBins = 5 * 16 * 8;
constexpr int kTileBins = 5 * 32 * 8;
// strip: tiles per work item (a multiple of 4; 0 picks one for the launch
// size with choose_strip3); max_groups > 0 caps the grid (persistent
// workgroups), 0 launches one workgroup per item.  A persistent grid needs
// wq: 9 zeroed device counters (work items taken per XCD, workgroups done),
// which every launch leaves zeroed again; launches sharing them must not
// overlap.  Without wq the launch is one workgroup per item.
// half: every slab's fields alternate in line parity (slabs_alternate), so
// the kernel that derives the rows ky >= 16 of each tile from the rows
// ky < 16 applies: in plain mode when g.thr_uniform >= 0 as well.  The
// features and apply-forward modes require it (they launch nothing
// without); the spectra and apply modes (full spectra in and out) ignore it.
constexpr int kWorkCounters = 9;
void launch_tpal3d(const Geom &g, const SlabJob *jobs, int n_jobs, const float *thr32, const double *thr64,
                   hipStream_t s, int mode = kTilePlain, void *spec = nullptr, const float *mask = nullptr,
                   int strip = 0, int max_groups = 0, unsigned *wq = nullptr, bool half = false);
// True when every slab's 8 fields alternate in line parity (host copies).
inline bool slabs_alternate(const SlabJob *jobs, int n_jobs)
{
    for (int i = 0; i < n_jobs; ++i)
        for (int z = 0; z + 1 < kZT; ++z)
            if (((jobs[i].offset[z] ^ jobs[i].offset[z + 1]) & 1) == 0) return false;
    return true;
}
int choose_strip3(const Geom &g, int n_jobs);
void launch_tpal2d(const Geom &g, const FieldJob2D *jobs, int n_jobs, int n_ty_max, const float *thr32,
                   const double *thr64, hipStream_t s);
// nchroma: the FrameJob.nchroma every job of the launch has (4, 2, 1, or 0
// for CHD_DEC_PAL_2D, which filters the composite)
// raw: with luma coring on (g.ynr > 0), raw_frame_floats(g) floats per job
// for the uncored planes (hipErrorInvalidValue without them); ignored otherwise.
hipError_t launch_demod(const Geom &g, const FrameJob *jobs, int n_jobs, const float *sine, const float *cosine,
                        const float *cfilt, int n_lblocks, hipStream_t s, int nchroma, float *raw);
inline int64_t raw_frame_floats_host(const Geom &g) { return 3 * (int64_t)(g.last - g.first + 1) * g.aw; }
// shift: samples dropped from the start of each field (black appended)
void launch_pack(const Geom &g, const uint16_t *src, int64_t src_field_elems, uint16_t *dst, int n_fields,
                 hipStream_t s, int shift = 0);
// Fields staged contiguously (src_field_elems apart), each packed into its
// own slot of a packed store: field k goes to dst + slot[k] * g.field_elems,
// shifted by shift[k] samples as launch_pack does.
constexpr int kPackListMax = 32;
struct PackList {
    int n;
    int32_t slot[kPackListMax];
    int32_t shift[kPackListMax];
};
void launch_pack_list(const Geom &g, const uint16_t *src, int64_t src_field_elems, uint16_t *dst,
                      const PackList &list, hipStream_t s);
// `bytes` from pinned host memory (its device address) into device memory,
// read by a kernel over the bus, so no DMA engine carries it.
void launch_copy_in(const void *src, void *dst, size_t bytes, hipStream_t s);
void launch_fill_border(const Geom &g, void *out, int64_t n_frames, hipStream_t s);
// the active lines of frame-line parity `parity` of one frame, as the zero
// Y/U/V no field was decoded into (frames whose fields share one offset)
void launch_fill_other_parity(const Geom &g, void *out, int parity, hipStream_t s);
void launch_chroma_to_planes(const Geom &g, const float *src, float *dst, int64_t plane_stride, int n_fields,
                             hipStream_t s);

}  // namespace tpal
