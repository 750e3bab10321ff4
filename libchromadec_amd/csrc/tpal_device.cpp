// Device-resident Transform-PAL decode: plans, tables, job lists, launches.
//
// Job construction keeps the reference's per-frame Z lattice
// (TransformPal3D::filterFields at transform_pal_3d.cpp:151 anchors its Z
// tiles at the start of the window chd_decode_frame loads, i.e. at absolute
// fields 2f-4 and 2f for frame f).  A slab of 8 fields starting at field 2k
// therefore serves frame k (its planes z = 0,1) and frame k+2 (z = 4,5); a
// run of F frames needs F+2 slabs instead of the 2F the reference computes.
#include <chromadec/decoder.h>
#include <chromadec/tpal_device.h>
#include <chromadec/tpal_mask.h>

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "mask_net_state.h"
#include "tpal_device_internal.h"
#include "tpal_internal.h"

using tpal::Geom;

namespace {

constexpr int kMaxRunsPerChunk = 16;

std::string hip_msg(const char *what, hipError_t e)
{
    return std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
}

double window(int n, int N) { return 0.5 - 0.5 * std::cos((2.0 * M_PI * (n + 0.5)) / N); }

tpal::Tables make_tables()
{
    tpal::Tables t;
    std::memset(&t, 0, sizeof(t));
    for (int z = 0; z < 8; ++z) { t.wz_d[z] = window(z, 8); t.wz[z] = (float)t.wz_d[z]; }
    for (int y = 0; y < 32; ++y) { t.wy_d[y] = window(y, 32); t.wy[y] = (float)t.wy_d[y]; }
    for (int x = 0; x < 16; ++x) { t.wx_d[x] = window(x, 16); t.wx[x] = (float)t.wx_d[x]; }
    for (int y = 0; y < 16; ++y) { t.wy2_d[y] = window(y, 16); t.wy2[y] = (float)t.wy2_d[y]; }
    for (int x = 0; x < 32; ++x) { t.wx2_d[x] = window(x, 32); t.wx2[x] = (float)t.wx2_d[x]; }
    // cos / sin (2 pi j / 32) from the first octant by the exact symmetries
    // of the circle: the quarter turns are exact and entry j + 16 is exactly
    // minus entry j, so the FP64 re-decision of a visit and of its image
    // half a period away in ky and kz gives bins of exactly opposite sign
    double c8[9];
    for (int r = 0; r <= 8; ++r) c8[r] = r == 0 ? 1.0 : r == 8 ? 0.0 : std::cos(2.0 * M_PI * r / 32.0);
    for (int j = 0; j < 32; ++j) {
        const int q = j / 8, r = j % 8;
        const double c = q == 0 ? c8[r] : q == 1 ? -c8[8 - r] : q == 2 ? -c8[r] : c8[8 - r];
        const double s = q == 0 ? c8[8 - r] : q == 1 ? c8[r] : q == 2 ? -c8[8 - r] : -c8[r];
        t.cos32_d[j] = c;
        t.sin32_d[j] = s;
    }
    return t;
}

std::mutex g_tables_mu;
std::vector<bool> g_tables_ready;

chd_status_t ensure_tables(int device)
{
    std::lock_guard<std::mutex> lk(g_tables_mu);
    if ((int)g_tables_ready.size() <= device) g_tables_ready.resize(device + 1, false);
    if (g_tables_ready[device]) return CHD_OK;
    const tpal::Tables t = make_tables();
    const hipError_t e = tpal::upload_tables(t);
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_tpal_plan_create: table upload", e));
    g_tables_ready[device] = true;
    return CHD_OK;
}

int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace

// ---------------------------------------------------------------------------

chd_status_t tpal_plan_create_impl(const chd_tpal_params_t *p, chd_tpal_plan **out, const char *who)
{
    const std::string W = who;
    if (p == nullptr || out == nullptr) return chd::fail(CHD_E_INVALID_ARG, W + ": null argument");
    if (p->kind != CHD_DEC_TRANSFORM_2D && p->kind != CHD_DEC_TRANSFORM_3D && p->kind != CHD_DEC_PAL_2D)
        return chd::fail(CHD_E_DECODER_UNKNOWN,
                         W + ": kind must be CHD_DEC_PAL_2D, CHD_DEC_TRANSFORM_2D or CHD_DEC_TRANSFORM_3D");
    if (p->pixel_format != CHD_PIXEL_YUV444PS && p->pixel_format != CHD_PIXEL_YUV444P16 &&
        p->pixel_format != CHD_PIXEL_GRAYS && p->pixel_format != CHD_PIXEL_GRAY16 &&
        p->pixel_format != CHD_PIXEL_RGB48 && p->pixel_format != CHD_PIXEL_RGBS)
        return chd::fail(CHD_E_UNSUPPORTED, W + ": pixel format not supported by the GPU decoder");
    if (p->field_width <= 0 || p->field_height <= 0 || p->active_video_start < 0 ||
        p->active_video_end > p->field_width || p->active_video_end <= p->active_video_start ||
        p->first_active_frame_line < 0 || p->last_active_frame_line < p->first_active_frame_line ||
        p->last_active_frame_line >= 2 * p->field_height - 1 || p->colour_burst_start < 0 ||
        p->colour_burst_end > p->field_width || p->colour_burst_end <= p->colour_burst_start ||
        p->sample_rate_hz <= 0.0 || p->white_16b_ire <= p->black_16b_ire)
        return chd::fail(CHD_E_INVALID_ARG, W + ": invalid video geometry");

    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
        return chd::fail(CHD_E_UNSUPPORTED, W + ": no HIP device available (the Transform PAL decoder runs on the GPU)");
    if (p->device < 0 || p->device >= ndev) return chd::fail(CHD_E_INVALID_ARG, W + ": device ordinal out of range");
    e = hipSetDevice(p->device);
    if (e != hipSuccess) return chd::fail(CHD_E_UNSUPPORTED, hip_msg((W + ": hipSetDevice").c_str(), e));

    auto plan = new chd_tpal_plan();
    plan->kind = p->kind;
    plan->device = p->device;
    Geom &g = plan->g;
    std::memset(&g, 0, sizeof(g));
    g.fw = p->field_width;
    g.fh = p->field_height;
    g.avs = p->active_video_start;
    g.ave = p->active_video_end;
    g.aw = g.ave - g.avs;
    g.first = p->first_active_frame_line;
    g.last = p->last_active_frame_line;
    g.cbs = p->colour_burst_start;
    g.cbe = p->colour_burst_end;
    g.black = (float)p->black_16b_ire;
    // packed field layout: tile origins (avs - 8 + 8k, avs - 16 + 16k) 8-aligned
    g.lmargin = 16 + ((8 - (g.avs % 8)) % 8);
    g.pitch = ((g.lmargin + g.fw + 32) + 63) / 64 * 64;
    g.field_elems = (int64_t)g.fh * g.pitch;
    g.cpitch = ((g.aw + 3 * tpal::kColOrigin) + 31) / 32 * 32;
    g.part_plane = (int64_t)g.fh * g.cpitch;
    g.n_ty3 = (g.last - (g.first - tpal::kYT / 2)) / (tpal::kYT / 2) + 1;
    g.n_tx3 = ceil_div(g.aw + tpal::kXT / 2, tpal::kXT / 2);
    g.n_strips3 = ceil_div(g.n_tx3, tpal::kStrip3);
    g.n_tx2 = ceil_div(g.aw + tpal::kXT2 / 2, tpal::kXT2 / 2);
    g.n_strips2 = ceil_div(g.n_tx2, tpal::kStrip2);
    plan->n_ty2_max = 0;
    plan->n_lblocks = 0;
    for (int off = 0; off < 2; ++off) {
        const int firstFL = (g.first + 1 - off) / 2;
        const int lastFL = ((g.last - off) / 2) + 1;
        const int nty = ceil_div(lastFL - (firstFL - tpal::kYT2 / 2), tpal::kYT2 / 2);
        plan->n_ty2_max = std::max(plan->n_ty2_max, nty);
        plan->n_lblocks = std::max(plan->n_lblocks, ceil_div(std::max(lastFL - firstFL, 0), tpal::kDemodLines));
    }
    g.burst_floor = (float)(130000.0 / 128);
    g.out_w = p->out_width;
    g.out_h = p->out_height;
    g.pad_left = p->pad_left;
    g.pad_top = p->pad_top;
    g.pixel_format = p->pixel_format;
    g.clamp_mode = p->clamp;
    if (g.out_w < g.aw + g.pad_left || g.out_h < (g.last - g.first + 1) + g.pad_top || g.pad_left < 0 || g.pad_top < 0) {
        delete plan;
        return chd::fail(CHD_E_INVALID_ARG, W + ": output frame smaller than the active picture");
    }
    const double range = (double)p->white_16b_ire - (double)p->black_16b_ire;
    g.ey_scale = (float)(1.0 / range);
    g.ecb_scale = (float)(p->ecb_from_u / range);
    g.ecr_scale = (float)(p->ecr_from_v / range);
    const double theta = (p->chroma_phase_deg * M_PI) / 180.0;
    g.chroma_gain_cos = (float)(std::cos(theta) * p->chroma_gain);
    g.chroma_gain_sin = (float)(std::sin(theta) * p->chroma_gain);
    // PalColour::doYNR: nr_y = yNRLevel * irescale, irescale the int32
    // quotient (white - black) / 100 (palcolour.cpp:414-416); off unless > 0
    g.ynr = 0.0f;
    if (p->luma_nr_level > 0.0)
        g.ynr = (float)(p->luma_nr_level * (double)((p->white_16b_ire - p->black_16b_ire) / 100));
    if (!(g.ynr > 0.0f)) g.ynr = 0.0f;
    {
        // Y'UV -> R'G'B' (color_conversion.cpp:33-67): the composite U/V
        // reductions follow from ecbFromU = 1 / (2 (1 - kb) uReduction)
        const double kr = p->kr > 0.0 ? p->kr : 0.299, kb = p->kb > 0.0 ? p->kb : 0.114;
        const double kg = 1.0 - kr - kb;
        const double ured = 1.0 / (2.0 * (1.0 - kb) * p->ecb_from_u);
        const double vred = 1.0 / (2.0 * (1.0 - kr) * p->ecr_from_v);
        g.rgb[0] = (float)(1.0 / vred);
        g.rgb[1] = (float)(-(kb / kg) / ured);
        g.rgb[2] = (float)(-(kr / kg) / vred);
        g.rgb[3] = (float)(1.0 / ured);
        // floatRGBBounds (output_writer.cpp:149-190)
        const float inf = INFINITY;
        for (int c = 0; c < 3; ++c) {
            g.rgb_lo[c] = -inf;
            g.rgb_hi[c] = inf;
        }
        if (p->clamp == CHD_CLAMP_LEGAL_RGB_SDR) {
            for (int c = 0; c < 3; ++c) {
                g.rgb_lo[c] = 0.0f;
                g.rgb_hi[c] = 1.0f;
            }
        } else if (p->clamp == CHD_CLAMP_LEGAL_RGB_HDR) {
            for (int c = 0; c < 3; ++c) g.rgb_lo[c] = 0.0f;
        } else if (p->clamp == CHD_CLAMP_LEGAL_YCBCR_BT601) {
            // the BT.601 legal Y'CbCr box carried through the inverse matrix
            const float eyLo = -15.0f / 219.0f, eyHi = 238.75f / 219.0f;
            const float ecLo = -127.0f / 224.0f, ecHi = 126.75f / 224.0f;
            const double rds = 2.0 * (1.0 - kr), bds = 2.0 * (1.0 - kb);
            const double gc = (kb * bds) / kg + (kr * rds) / kg;
            g.rgb_lo[0] = (float)(eyLo + rds * ecLo);
            g.rgb_hi[0] = (float)(eyHi + rds * ecHi);
            g.rgb_lo[1] = (float)(eyLo - gc * ecHi);
            g.rgb_hi[1] = (float)(eyHi - gc * ecLo);
            g.rgb_lo[2] = (float)(eyLo + bds * ecLo);
            g.rgb_hi[2] = (float)(eyHi + bds * ecHi);
        }
    }

    // PalColour::buildLookUpTables (palcolour.cpp:150-277)
    const double bw = 1100000.0 / 0.93;
    const double ca = 0.5 * p->sample_rate_hz / bw;
    const int fs = std::max(1, (int)ca);
    if (fs > tpal::kMaxFilter) {
        delete plan;
        return chd::fail(CHD_E_UNSUPPORTED, W + ": sample rate needs a chroma filter wider than the GPU kernel supports");
    }
    g.filter_size = fs;
    // chroma taps cfilt and luma taps yfilt (the luma filter has the same
    // width: ya = ca), interleaved per tap with stride kFiltStride
    constexpr int FS = tpal::kFiltStride;
    std::vector<float> cfilt((tpal::kMaxFilter + 1) * FS, 0.0f);
    {
        std::vector<double> cf((fs + 1) * 4, 0.0), yf((fs + 1) * 2, 0.0);
        double cdiv = 0, ydiv = 0;
        for (int f = 0; f <= fs; ++f) {
            const double fc = std::min(ca, (double)f);
            const double ff = std::min(ca, std::sqrt((double)f * f + 2 * 2));
            const double fff = std::min(ca, std::sqrt((double)f * f + 4 * 4));
            const double ffff = std::min(ca, std::sqrt((double)f * f + 6 * 6));
            const int d = (f == 0) ? 2 : 1;
            cf[f * 4 + 0] = (1 + std::cos(M_PI * fc / ca)) / d;
            cf[f * 4 + 2] = (1 + std::cos(M_PI * ff / ca)) / d;
            cf[f * 4 + 1] = (1 + std::cos(M_PI * fff / ca)) / d;
            cf[f * 4 + 3] = (1 + std::cos(M_PI * ffff / ca)) / d;
            cdiv += 2 * (cf[f * 4 + 0] + 2 * cf[f * 4 + 2] + 2 * cf[f * 4 + 1] + 2 * cf[f * 4 + 3]);
            // Y filter: lines 0 and +-2 only, the outer taps weighted 0.2
            yf[f * 2 + 0] = (1 + std::cos(M_PI * fc / ca)) / d;
            yf[f * 2 + 1] = 0.2 * (1 + std::cos(M_PI * fff / ca)) / d;
            ydiv += 2 * (yf[f * 2 + 0] + 2 * yf[f * 2 + 1]);
        }
        for (int f = 0; f <= fs; ++f) {
            for (int j = 0; j < 4; ++j) cfilt[f * FS + j] = (float)(cf[f * 4 + j] / cdiv);
            for (int j = 0; j < 2; ++j) cfilt[f * FS + 4 + j] = (float)(yf[f * 2 + j] / ydiv);
        }
    }
    std::vector<float> sine(g.fw), cosine(g.fw);
    // At sampleRate = 4 fSC the carrier samples are exactly 0 and +-1 (the
    // reference's FP64 sin/cos leave residues near 1e-16 where they vanish),
    // which the quadrature demodulator exploits.
    const bool quad = std::fabs(p->fsc_hz / p->sample_rate_hz - 0.25) < 1e-12;
    for (int i = 0; i < g.fw; ++i) {
        const double rad = 2 * M_PI * i * p->fsc_hz / p->sample_rate_hz;
        double sv = std::sin(rad), cv = std::cos(rad);
        if (quad) {
            sv = std::round(sv);
            cv = std::round(cv);
        }
        // PAL-M swaps the references (palcolour.cpp:158-174)
        sine[i] = (float)(g.fh != 263 ? sv : cv);
        cosine[i] = (float)(g.fh != 263 ? cv : sv);
    }
    g.quad = (quad && fs <= 8 && g.avs >= fs && g.ave + fs + 1 <= g.fw) ? 1 : 0;
    g.spar = (g.fw > 1 && sine[1] != 0.0f) ? 1 : 0;

    // thresholds (transform_pal.cpp:53-77): squared, z/y/x order
    const int nthr = (p->kind == CHD_DEC_TRANSFORM_3D) ? 768 : 80;
    std::vector<double> thr64(nthr);
    std::vector<float> thr32(nthr);
    bool uniform = true;
    for (int i = 0; i < nthr; ++i) {
        const double t = p->thresholds ? p->thresholds[i] : p->threshold;
        thr64[i] = t * t;
        thr32[i] = (float)thr64[i];
        uniform = uniform && thr32[i] == thr32[0];
    }
    g.thr_uniform = uniform ? thr32[0] : -1.0f;
    g.thr_uniform_d = uniform ? thr64[0] : -1.0;

    chd_status_t st = ensure_tables(p->device);
    if (st != CHD_OK) { delete plan; return st; }

    plan->max_frames = p->max_batch_frames > 0 ? p->max_batch_frames : 64;
    // workspace units: 3D slabs (frames + 2 per run), 2D fields; PAL_2D
    // decodes straight from the fields but keeps field units for the
    // separated-chroma entry point
    if (p->kind == CHD_DEC_TRANSFORM_3D) {
        plan->max_units = plan->max_frames + 2 * std::min(kMaxRunsPerChunk, plan->max_frames);
        plan->unit_floats = 8 * (size_t)g.part_plane;
    } else {
        plan->max_units = 2 * plan->max_frames;
        plan->unit_floats = 2 * (size_t)g.part_plane;
    }
    auto dmalloc = [&](void **ptr, size_t bytes) -> bool {
        const hipError_t r = hipMalloc(ptr, bytes);
        if (r != hipSuccess) { *ptr = nullptr; return false; }
        return true;
    };
    bool ok = true;
    ok = ok && dmalloc((void **)&plan->d_thr32, nthr * sizeof(float));
    ok = ok && dmalloc((void **)&plan->d_thr64, nthr * sizeof(double));
    ok = ok && dmalloc((void **)&plan->d_sine, g.fw * sizeof(float));
    ok = ok && dmalloc((void **)&plan->d_cosine, g.fw * sizeof(float));
    ok = ok && dmalloc((void **)&plan->d_cfilt, cfilt.size() * sizeof(float));
    ok = ok && dmalloc((void **)&plan->d_partials, plan->max_units * plan->unit_floats * sizeof(float));
    ok = ok && dmalloc((void **)&plan->d_slab_jobs, plan->max_units * sizeof(tpal::SlabJob));
    ok = ok && dmalloc((void **)&plan->d_field_jobs, plan->max_units * sizeof(tpal::FieldJob2D));
    ok = ok && dmalloc((void **)&plan->d_frame_jobs, plan->max_frames * sizeof(tpal::FrameJob));
    if (g.ynr > 0.0f)
        ok = ok && dmalloc((void **)&plan->d_raw, plan->max_frames * tpal::raw_frame_floats_host(g) * sizeof(float));
    ok = ok && hipHostMalloc((void **)&plan->h_slab_jobs, plan->max_units * sizeof(tpal::SlabJob), 0) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&plan->h_field_jobs, plan->max_units * sizeof(tpal::FieldJob2D), 0) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&plan->h_frame_jobs, plan->max_frames * sizeof(tpal::FrameJob), 0) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&plan->staging_done, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&plan->call_done, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        chd_tpal_plan_free(plan);
        return chd::fail(CHD_E_OOM, W + ": device allocation failed");
    }
    ok = ok && hipMemcpy(plan->d_thr32, thr32.data(), nthr * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(plan->d_thr64, thr64.data(), nthr * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(plan->d_sine, sine.data(), g.fw * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(plan->d_cosine, cosine.data(), g.fw * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(plan->d_cfilt, cfilt.data(), cfilt.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        chd_tpal_plan_free(plan);
        return chd::fail(CHD_E_INTERNAL, W + ": device upload failed");
    }
    *out = plan;
    return CHD_OK;
}

// Kernel timing.  The events of the recorded chunks are summed into the
// totals when they are read, so that recording never waits for the GPU.
static void timing_reset(chd_tpal_plan *plan)
{
    plan->t_pending = 0;
    plan->last_tile_ms = plan->last_demod_ms = 0;
    plan->n_tile_launches = plan->n_demod_launches = 0;
}

// A decode call starts new totals unless they run over every call (mode 2).
static void timing_begin(chd_tpal_plan *plan)
{
    if (plan->timing != 2) timing_reset(plan);
}

// Adds the recorded chunks' tile and demod times to the totals (waits for
// the last recorded chunk).
static void timing_resolve(chd_tpal_plan *plan)
{
    if (plan->t_pending == 0) return;
    hipEventSynchronize(plan->t_events[3 * plan->t_pending - 1]);
    for (int i = 0; i < plan->t_pending; ++i) {
        float a = 0, b = 0;
        hipEventElapsedTime(&a, plan->t_events[3 * i], plan->t_events[3 * i + 1]);
        hipEventElapsedTime(&b, plan->t_events[3 * i + 1], plan->t_events[3 * i + 2]);
        plan->last_tile_ms += a;
        plan->last_demod_ms += b;
        plan->n_tile_launches += 1;
        plan->n_demod_launches += 1;
    }
    plan->t_pending = 0;
}

// The three events of the next recorded chunk, or nullptr when an event
// cannot be created (the chunk then goes untimed).  At most 1024 chunks are
// held before they are summed.
static hipEvent_t *timing_slot(chd_tpal_plan *plan)
{
    constexpr int kMaxTimedChunks = 1024;
    if (plan->t_pending == kMaxTimedChunks) timing_resolve(plan);
    const size_t need = 3 * (size_t)(plan->t_pending + 1);
    while (plan->t_events.size() < need) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        plan->t_events.push_back(e);
    }
    return &plan->t_events[3 * (size_t)plan->t_pending++];
}

// Enqueue one chunk: copy job lists, launch the tile and demod kernels.
// Partial chroma planes per frame job (FrameJob.nchroma) for the plan's kind.
static int plan_nchroma(const chd_tpal_plan *plan)
{
    return plan->kind == CHD_DEC_TRANSFORM_3D ? 4 : plan->kind == CHD_DEC_TRANSFORM_2D ? 2 : 0;
}

static chd_status_t run_chunk(chd_tpal_plan *plan, const TpalWorkspace &ws, int n_units, int n_frames, hipStream_t s,
                              const std::string &who)
{
    const Geom &g = plan->g;
    hipError_t e = hipSuccess;
    // A repeated decode of the same frames into the same output (a benchmark
    // loop, a re-render) builds the same job lists: when the plan's own job
    // buffers already hold them, the two uploads (and the wait for them
    // before the tile launch) are skipped.  Calls are ordered on the plan
    // (call_begin), so the buffers are not rewritten under a running call.
    const bool is3d = plan->kind == CHD_DEC_TRANSFORM_3D, is2d = plan->kind == CHD_DEC_TRANSFORM_2D;
    const size_t unit_bytes = is3d ? n_units * sizeof(tpal::SlabJob) : is2d ? n_units * sizeof(tpal::FieldJob2D) : 0;
    const unsigned char *unit_src = is3d ? reinterpret_cast<const unsigned char *>(plan->h_slab_jobs)
                                         : reinterpret_cast<const unsigned char *>(plan->h_field_jobs);
    const unsigned char *frame_src = reinterpret_cast<const unsigned char *>(plan->h_frame_jobs);
    const size_t frame_job_bytes = n_frames * sizeof(tpal::FrameJob);
    const bool own = ws.frame_jobs == plan->d_frame_jobs;
    const bool same = own && plan->up_valid && plan->up_shadow.size() == unit_bytes + frame_job_bytes &&
                      std::memcmp(plan->up_shadow.data(), unit_src, unit_bytes) == 0 &&
                      std::memcmp(plan->up_shadow.data() + unit_bytes, frame_src, frame_job_bytes) == 0;
    if (!same) {
        if (is3d) {
            e = hipMemcpyAsync(ws.slab_jobs, plan->h_slab_jobs, unit_bytes, hipMemcpyHostToDevice, s);
        } else if (is2d) {
            e = hipMemcpyAsync(ws.field_jobs, plan->h_field_jobs, unit_bytes, hipMemcpyHostToDevice, s);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(ws.frame_jobs, frame_src, frame_job_bytes, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(plan->staging_done, s);
        if (e != hipSuccess) {
            plan->up_valid = false;
            return chd::fail(CHD_E_INTERNAL, hip_msg((who + ": job upload").c_str(), e));
        }
        if (own) {
            plan->up_shadow.assign(unit_src, unit_src + unit_bytes);
            plan->up_shadow.insert(plan->up_shadow.end(), frame_src, frame_src + frame_job_bytes);
            plan->up_valid = true;
        }
    }
    const bool border = (g.out_w != g.aw || g.out_h != g.last - g.first + 1);
    if (border) {
        for (int i = 0; i < n_frames; ++i) tpal::launch_fill_border(g, plan->h_frame_jobs[i].out, 1, s);
    }
    hipEvent_t *tev = plan->timing ? timing_slot(plan) : nullptr;
    if (tev) hipEventRecord(tev[0], s);
    if (plan->kind == CHD_DEC_TRANSFORM_3D && plan->net != nullptr) {
        // learned mask: tile spectra -> network -> masked inverse
        chd_mask_net *net = plan->net;
        const int64_t tiles = (int64_t)n_units * g.n_ty3 * g.n_tx3;
        // alternating parity: magnitude features of rows ky < 16, paired
        // inference, the half-form inverse after the tile transform computed
        // again from the composite; otherwise full spectra stored and read
        // back
        const bool paired = tpal::slabs_alternate(plan->h_slab_jobs, n_units);
        tpal::launch_tpal3d(g, ws.slab_jobs, n_units, plan->d_thr32, plan->d_thr64, s,
                            paired ? tpal::kTileFeatures : tpal::kTileSpectra, plan->d_spec, nullptr, 0, 0, nullptr,
                            paired);
        mnet::launch_infer(plan->d_spec, tiles, net->theta, net->f, plan->d_mask, net->cus, s, paired);
        tpal::launch_tpal3d(g, ws.slab_jobs, n_units, plan->d_thr32, plan->d_thr64, s,
                            paired ? tpal::kTileApplyFwd : tpal::kTileApply, plan->d_spec, plan->d_mask, 0, 0, nullptr,
                            paired);
    } else if (plan->kind == CHD_DEC_TRANSFORM_3D)
        tpal::launch_tpal3d(g, ws.slab_jobs, n_units, plan->d_thr32, plan->d_thr64, s, tpal::kTilePlain, nullptr,
                            nullptr, 0, 0, nullptr, tpal::slabs_alternate(plan->h_slab_jobs, n_units));
    else if (plan->kind == CHD_DEC_TRANSFORM_2D)
        tpal::launch_tpal2d(g, ws.field_jobs, n_units, plan->n_ty2_max, plan->d_thr32, plan->d_thr64, s);
    if (tev) hipEventRecord(tev[1], s);
    e = tpal::launch_demod(g, ws.frame_jobs, n_frames, plan->d_sine, plan->d_cosine, plan->d_cfilt,
                           plan->n_lblocks, s, plan_nchroma(plan), ws.raw);
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg((who + ": demod launch").c_str(), e));
    for (int i = 0; i < n_frames; ++i) {
        const tpal::FrameJob &fj = plan->h_frame_jobs[i];
        if (fj.offset[0] == fj.offset[1]) tpal::launch_fill_other_parity(g, fj.out, fj.offset[0] ^ 1, s);
    }
    if (tev) hipEventRecord(tev[2], s);
    e = hipGetLastError();
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg((who + ": kernel launch").c_str(), e));
    return CHD_OK;
}

static void wait_staging(chd_tpal_plan *plan)
{
    if (plan->staging_pending) {
        hipEventSynchronize(plan->staging_done);
        plan->staging_pending = false;
    }
}

// The workspace a plan owns.
static TpalWorkspace plan_ws(const chd_tpal_plan *plan)
{
    TpalWorkspace ws;
    ws.partials = plan->d_partials;
    ws.max_units = plan->max_units;
    ws.max_frames = plan->max_frames;
    ws.slab_jobs = plan->d_slab_jobs;
    ws.field_jobs = plan->d_field_jobs;
    ws.frame_jobs = plan->d_frame_jobs;
    ws.raw = plan->d_raw;
    return ws;
}

// A call enqueued on a different stream than the previous one must not
// overwrite partials or raw planes that the previous call's kernels still
// read: order it after that call's last kernel.
static void call_begin(chd_tpal_plan *plan, hipStream_t s)
{
    if (plan->call_pending && plan->last_stream != s) hipStreamWaitEvent(s, plan->call_done, 0);
}

static void call_end(chd_tpal_plan *plan, hipStream_t s)
{
    if (hipEventRecord(plan->call_done, s) == hipSuccess) {
        plan->call_pending = true;
        plan->last_stream = s;
    }
}

static int64_t frame_bytes(const chd_tpal_plan *plan)
{
    const Geom &g = plan->g;
    const int planes = (g.pixel_format == CHD_PIXEL_GRAY16 || g.pixel_format == CHD_PIXEL_GRAYS) ? 1 : 3;
    const int es = (g.pixel_format == CHD_PIXEL_YUV444PS || g.pixel_format == CHD_PIXEL_GRAYS ||
                    g.pixel_format == CHD_PIXEL_RGBS) ? 4 : 2;
    return (int64_t)planes * g.out_w * g.out_h * es;
}

chd_status_t tpal_decode_windows(chd_tpal_plan *plan, const TpalWindow *w, int n, hipStream_t s)
{
    // Each window is decoded on its own lattice (two slabs / two fields).
    const Geom &g = plan->g;
    hipSetDevice(plan->device);
    timing_begin(plan);
    const TpalWorkspace ws = plan_ws(plan);
    call_begin(plan, s);
    const int per_chunk = std::max(1, std::min(plan->max_frames, plan->max_units / 2));
    for (int base = 0; base < n; base += per_chunk) {
        const int cnt = std::min(per_chunk, n - base);
        wait_staging(plan);
        for (int i = 0; i < cnt; ++i) {
            const TpalWindow &win = w[base + i];
            tpal::FrameJob &fj = plan->h_frame_jobs[i];
            std::memset(&fj, 0, sizeof(fj));
            fj.out = win.out;
            if (plan->kind == CHD_DEC_TRANSFORM_3D) {
                // window fields 0..13, start 4: slabs at window offsets 0 and 4
                for (int sl = 0; sl < 2; ++sl) {
                    tpal::SlabJob &sj = plan->h_slab_jobs[2 * i + sl];
                    for (int z = 0; z < 8; ++z) {
                        sj.field[z] = win.field[4 * sl + z];
                        sj.offset[z] = win.offset[4 * sl + z];
                    }
                    sj.partial = plan->d_partials + (size_t)(2 * i + sl) * plan->unit_floats;
                }
                for (int fsel = 0; fsel < 2; ++fsel) {
                    fj.field[fsel] = win.field[4 + fsel];
                    fj.offset[fsel] = win.offset[4 + fsel];
                    const float *a = plan->h_slab_jobs[2 * i].partial + (2 + fsel) * 2 * g.part_plane;
                    const float *b = plan->h_slab_jobs[2 * i + 1].partial + fsel * 2 * g.part_plane;
                    fj.chroma[fsel][0] = a;
                    fj.chroma[fsel][1] = a + g.part_plane;
                    fj.chroma[fsel][2] = b;
                    fj.chroma[fsel][3] = b + g.part_plane;
                }
                fj.nchroma = 4;
            } else if (plan->kind == CHD_DEC_TRANSFORM_2D) {
                for (int fsel = 0; fsel < 2; ++fsel) {
                    tpal::FieldJob2D &jj = plan->h_field_jobs[2 * i + fsel];
                    jj.field = win.field[fsel];
                    jj.offset = win.offset[fsel];
                    jj.partial = plan->d_partials + (size_t)(2 * i + fsel) * plan->unit_floats;
                    fj.field[fsel] = win.field[fsel];
                    fj.offset[fsel] = win.offset[fsel];
                    fj.chroma[fsel][0] = jj.partial;
                    fj.chroma[fsel][1] = jj.partial + g.part_plane;
                }
                fj.nchroma = 2;
            } else {
                for (int fsel = 0; fsel < 2; ++fsel) {
                    fj.field[fsel] = win.field[fsel];
                    fj.offset[fsel] = win.offset[fsel];
                }
                fj.nchroma = 0;
            }
        }
        chd_status_t st = run_chunk(plan, ws, 2 * cnt, cnt, s, "chd_decode_frame");
        plan->staging_pending = true;
        if (st != CHD_OK) return st;
    }
    call_end(plan, s);
    return CHD_OK;
}

// ---------------------------------------------------------------------------

extern "C" {

chd_status_t chd_tpal_plan_create(const chd_tpal_params_t *params, chd_tpal_plan_t **out)
{
    return tpal_plan_create_impl(params, out, "chd_tpal_plan_create");
}

void chd_tpal_plan_free(chd_tpal_plan_t *plan)
{
    if (plan == nullptr) return;
    hipSetDevice(plan->device);
    if (plan->staging_pending) hipEventSynchronize(plan->staging_done);
    hipDeviceSynchronize();
    hipFree(plan->d_thr32);
    hipFree(plan->d_thr64);
    hipFree(plan->d_sine);
    hipFree(plan->d_cosine);
    hipFree(plan->d_cfilt);
    hipFree(plan->d_partials);
    hipFree(plan->d_slab_jobs);
    hipFree(plan->d_field_jobs);
    hipFree(plan->d_frame_jobs);
    if (plan->d_raw) hipFree(plan->d_raw);
    if (plan->d_spec) hipFree(plan->d_spec);
    if (plan->d_mask) hipFree(plan->d_mask);
    if (plan->h_slab_jobs) hipHostFree(plan->h_slab_jobs);
    if (plan->h_field_jobs) hipHostFree(plan->h_field_jobs);
    if (plan->h_frame_jobs) hipHostFree(plan->h_frame_jobs);
    if (plan->staging_done) hipEventDestroy(plan->staging_done);
    if (plan->call_done) hipEventDestroy(plan->call_done);
    for (hipEvent_t e : plan->t_events) hipEventDestroy(e);
    delete plan;
}

size_t chd_tpal_packed_field_elems(const chd_tpal_plan_t *plan)
{
    return plan ? (size_t)plan->g.field_elems : 0;
}

size_t chd_tpal_frame_bytes(const chd_tpal_plan_t *plan) { return plan ? (size_t)frame_bytes(plan) : 0; }

chd_status_t chd_tpal_pack_fields(chd_tpal_plan_t *plan, const uint16_t *d_src, int64_t n_fields,
                                  uint16_t *d_packed, void *stream)
{
    if (plan == nullptr || (n_fields > 0 && (d_src == nullptr || d_packed == nullptr)) || n_fields < 0)
        return chd::fail(CHD_E_INVALID_ARG, "chd_tpal_pack_fields: null argument");
    hipSetDevice(plan->device);
    tpal::launch_pack(plan->g, d_src, (int64_t)plan->g.fw * plan->g.fh, d_packed, (int)n_fields,
                      (hipStream_t)stream);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_tpal_pack_fields", e));
    return CHD_OK;
}

}  // extern "C"

chd_status_t tpal_check_runs(const chd_tpal_plan *plan, const chd_tpal_sequence_t *seqs, int n_seqs, void *d_out,
                             const char *who)
{
    const std::string W = who;
    if (plan == nullptr || (n_seqs > 0 && (seqs == nullptr || d_out == nullptr)) || n_seqs < 0)
        return chd::fail(CHD_E_INVALID_ARG, W + ": null argument");
    for (int i = 0; i < n_seqs; ++i) {
        const chd_tpal_sequence_t &q = seqs[i];
        if (q.count < 0 || q.first_frame < 0 || q.first_frame + q.count > q.n_frames ||
            (q.count > 0 && q.d_fields == nullptr))
            return chd::fail(CHD_E_OUT_OF_RANGE, W + ": frame run outside its sequence");
    }
    return CHD_OK;
}

chd_status_t tpal_decode_runs(chd_tpal_plan *plan, const TpalWorkspace &ws, const chd_tpal_sequence_t *seqs,
                              int n_seqs, void *d_out, hipStream_t s, const char *who)
{
    hipSetDevice(plan->device);
    const Geom &g = plan->g;
    const int64_t fb = frame_bytes(plan);
    const bool is3d = plan->kind == CHD_DEC_TRANSFORM_3D;
    timing_begin(plan);

    int units = 0, frames = 0, runs = 0;
    int64_t out_index = 0;
    auto flush = [&]() -> chd_status_t {
        if (frames == 0) return CHD_OK;
        chd_status_t st = run_chunk(plan, ws, units, frames, s, who);
        plan->staging_pending = true;
        units = frames = runs = 0;
        return st;
    };
    auto field_ptr = [&](const chd_tpal_sequence_t &q, int64_t fi) -> const uint16_t * {
        return (fi >= 0 && fi < 2 * q.n_frames) ? q.d_fields + fi * g.field_elems : nullptr;
    };
    wait_staging(plan);
    call_begin(plan, s);
    for (int si = 0; si < n_seqs; ++si) {
        const chd_tpal_sequence_t &q = seqs[si];
        int64_t f = q.first_frame;
        const int64_t fend = q.first_frame + q.count;
        while (f < fend) {
            // frames this run can take in the current chunk
            const int extra = is3d ? 2 : 0;
            const int unit_per_frame = is3d ? 1 : 2;
            int room_units = ws.max_units - units - extra;
            int room = std::min(ws.max_frames - frames, room_units / unit_per_frame);
            if (runs >= kMaxRunsPerChunk || room <= 0) {
                if (frames == 0) return chd::fail(CHD_E_INTERNAL, std::string(who) + ": workspace too small");
                chd_status_t st = flush();
                if (st != CHD_OK) return st;
                wait_staging(plan);
                continue;
            }
            const int cnt = (int)std::min<int64_t>(room, fend - f);
            const int unit0 = units;
            if (is3d) {
                // slabs k = f-2 .. f+cnt-1
                for (int k = 0; k < cnt + 2; ++k) {
                    tpal::SlabJob &sj = plan->h_slab_jobs[units + k];
                    const int64_t slab = f - 2 + k;
                    for (int z = 0; z < 8; ++z) {
                        sj.field[z] = field_ptr(q, 2 * slab + z);
                        sj.offset[z] = z & 1;
                    }
                    sj.partial = ws.partials + (size_t)(units + k) * plan->unit_floats;
                }
                units += cnt + 2;
            } else if (plan->kind == CHD_DEC_TRANSFORM_2D) {
                for (int k = 0; k < 2 * cnt; ++k) {
                    tpal::FieldJob2D &jj = plan->h_field_jobs[units + k];
                    jj.field = field_ptr(q, 2 * f + k);
                    jj.offset = k & 1;
                    jj.partial = ws.partials + (size_t)(units + k) * plan->unit_floats;
                }
                units += 2 * cnt;
            } else {
                units += 2 * cnt;   // PAL_2D: no transform work, the budget stays per field
            }
            for (int i = 0; i < cnt; ++i) {
                tpal::FrameJob &fj = plan->h_frame_jobs[frames + i];
                std::memset(&fj, 0, sizeof(fj));
                fj.out = (char *)d_out + (out_index + i) * fb;
                for (int fsel = 0; fsel < 2; ++fsel) {
                    fj.field[fsel] = field_ptr(q, 2 * (f + i) + fsel);
                    fj.offset[fsel] = fsel;
                    if (is3d) {
                        // slab f-2 (its z = 4,5) then slab f (z = 0,1)
                        const float *a = plan->h_slab_jobs[unit0 + i].partial + (2 + fsel) * 2 * g.part_plane;
                        const float *b = plan->h_slab_jobs[unit0 + i + 2].partial + fsel * 2 * g.part_plane;
                        fj.chroma[fsel][0] = a;
                        fj.chroma[fsel][1] = a + g.part_plane;
                        fj.chroma[fsel][2] = b;
                        fj.chroma[fsel][3] = b + g.part_plane;
                    } else if (plan->kind == CHD_DEC_TRANSFORM_2D) {
                        const float *a = plan->h_field_jobs[unit0 + 2 * i + fsel].partial;
                        fj.chroma[fsel][0] = a;
                        fj.chroma[fsel][1] = a + g.part_plane;
                    }
                }
                fj.nchroma = is3d ? 4 : (plan->kind == CHD_DEC_TRANSFORM_2D ? 2 : 0);
            }
            frames += cnt;
            out_index += cnt;
            runs += 1;
            f += cnt;
        }
    }
    const chd_status_t st = flush();
    if (st == CHD_OK) call_end(plan, s);
    return st;
}

chd_status_t chd_tpal_decode(chd_tpal_plan_t *plan, const chd_tpal_sequence_t *seqs, int32_t n_seqs, void *d_out,
                             void *stream)
{
    const chd_status_t st = tpal_check_runs(plan, seqs, n_seqs, d_out, "chd_tpal_decode");
    if (st != CHD_OK) return st;
    return tpal_decode_runs(plan, plan_ws(plan), seqs, n_seqs, d_out, (hipStream_t)stream, "chd_tpal_decode");
}

extern "C" {

int64_t chd_tpal_tiles_per_slab(const chd_tpal_plan_t *plan)
{
    return (plan && plan->kind == CHD_DEC_TRANSFORM_3D) ? (int64_t)plan->g.n_ty3 * plan->g.n_tx3 : 0;
}

// Spectra and masks of a whole chunk (max_units slabs) for the learned mask.
static chd_status_t ensure_spectra(chd_tpal_plan *plan, const char *who)
{
    if (plan->d_spec != nullptr) return CHD_OK;
    const size_t tiles = (size_t)plan->max_units * plan->g.n_ty3 * plan->g.n_tx3;
    hipError_t e = hipMalloc(&plan->d_spec, tiles * tpal::kTileBins * 2 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&plan->d_mask, tiles * mnet::kP * sizeof(float));
    if (e != hipSuccess) {
        hipFree(plan->d_spec);
        plan->d_spec = nullptr;
        plan->d_mask = nullptr;
        return chd::fail(CHD_E_OOM, hip_msg(who, e));
    }
    return CHD_OK;
}

chd_status_t chd_tpal_set_mask_net(chd_tpal_plan_t *plan, chd_mask_net_t *net)
{
    if (plan == nullptr) return chd::fail(CHD_E_INVALID_ARG, "chd_tpal_set_mask_net: null plan");
    if (net != nullptr && plan->kind != CHD_DEC_TRANSFORM_3D)
        return chd::fail(CHD_E_UNSUPPORTED, "chd_tpal_set_mask_net: the learned mask is for CHD_DEC_TRANSFORM_3D");
    if (net != nullptr && net->device != plan->device)
        return chd::fail(CHD_E_INVALID_ARG, "chd_tpal_set_mask_net: network and plan are on different devices");
    hipSetDevice(plan->device);
    wait_staging(plan);
    if (net != nullptr) {
        chd_status_t st = ensure_spectra(plan, "chd_tpal_set_mask_net");
        if (st != CHD_OK) return st;
    }
    plan->net = net;
    return CHD_OK;
}

chd_status_t chd_tpal_slab_spectra(chd_tpal_plan_t *plan, const uint16_t *d_fields, int64_t n_fields,
                                   int64_t first_slab, int32_t n_slabs, void *d_spectra, void *stream)
{
    if (plan == nullptr || n_fields < 0 || n_slabs < 0 || (n_slabs > 0 && d_spectra == nullptr) ||
        (n_fields > 0 && d_fields == nullptr))
        return chd::fail(CHD_E_INVALID_ARG, "chd_tpal_slab_spectra: null argument");
    if (plan->kind != CHD_DEC_TRANSFORM_3D)
        return chd::fail(CHD_E_UNSUPPORTED, "chd_tpal_slab_spectra: 3D plans only");
    hipSetDevice(plan->device);
    const hipStream_t s = (hipStream_t)stream;
    const Geom &g = plan->g;
    const int64_t tps = (int64_t)g.n_ty3 * g.n_tx3;
    call_begin(plan, s);
    for (int base = 0; base < n_slabs; base += plan->max_units) {
        const int cnt = std::min(plan->max_units, n_slabs - base);
        wait_staging(plan);
        for (int k = 0; k < cnt; ++k) {
            tpal::SlabJob &sj = plan->h_slab_jobs[k];
            const int64_t f0 = 2 * (first_slab + base + k);
            for (int z = 0; z < 8; ++z) {
                const int64_t fi = f0 + z;
                sj.field[z] = (fi >= 0 && fi < n_fields) ? d_fields + fi * g.field_elems : nullptr;
                sj.offset[z] = z & 1;
            }
            sj.partial = nullptr;
        }
        plan->up_valid = false;   // the decode's job buffers are rewritten
        hipError_t e = hipMemcpyAsync(plan->d_slab_jobs, plan->h_slab_jobs, cnt * sizeof(tpal::SlabJob),
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(plan->staging_done, s);
        if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_tpal_slab_spectra: job upload", e));
        plan->staging_pending = true;
        void *dst = (char *)d_spectra + (size_t)base * tps * tpal::kTileBins * 2 * sizeof(float);
        tpal::launch_tpal3d(g, plan->d_slab_jobs, cnt, plan->d_thr32, plan->d_thr64, s, tpal::kTileSpectra, dst,
                            nullptr);
        e = hipGetLastError();
        if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_tpal_slab_spectra: launch", e));
    }
    call_end(plan, s);
    return CHD_OK;
}

chd_status_t chd_tpal_decode_separated(chd_tpal_plan_t *plan, const uint16_t *d_fields, const float *d_chroma,
                                       int64_t n_frames, void *d_out, void *stream)
{
    if (plan == nullptr || n_frames < 0 || (n_frames > 0 && (d_fields == nullptr || d_chroma == nullptr ||
                                                            d_out == nullptr)))
        return chd::fail(CHD_E_INVALID_ARG, "chd_tpal_decode_separated: null argument");
    hipSetDevice(plan->device);
    const hipStream_t s = (hipStream_t)stream;
    const Geom &g = plan->g;
    const int64_t fb = frame_bytes(plan);
    const int per_chunk = std::max(1, std::min(plan->max_frames, plan->max_units / 2));
    const bool border = (g.out_w != g.aw || g.out_h != g.last - g.first + 1);
    call_begin(plan, s);
    for (int64_t base = 0; base < n_frames; base += per_chunk) {
        const int cnt = (int)std::min<int64_t>(per_chunk, n_frames - base);
        wait_staging(plan);
        tpal::launch_chroma_to_planes(g, d_chroma + base * 2 * (int64_t)g.fw * g.fh, plan->d_partials,
                                      (int64_t)plan->unit_floats, 2 * cnt, s);
        for (int i = 0; i < cnt; ++i) {
            tpal::FrameJob &fj = plan->h_frame_jobs[i];
            std::memset(&fj, 0, sizeof(fj));
            fj.out = (char *)d_out + (base + i) * fb;
            for (int fsel = 0; fsel < 2; ++fsel) {
                fj.field[fsel] = d_fields + (2 * (base + i) + fsel) * g.field_elems;
                fj.offset[fsel] = fsel;
                fj.chroma[fsel][0] = plan->d_partials + (size_t)(2 * i + fsel) * plan->unit_floats;
            }
            fj.nchroma = 1;
            if (border) tpal::launch_fill_border(g, fj.out, 1, s);
        }
        plan->up_valid = false;   // the decode's job buffers are rewritten
        hipError_t e = hipMemcpyAsync(plan->d_frame_jobs, plan->h_frame_jobs, cnt * sizeof(tpal::FrameJob),
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(plan->staging_done, s);
        if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_tpal_decode_separated: job upload", e));
        plan->staging_pending = true;
        e = tpal::launch_demod(g, plan->d_frame_jobs, cnt, plan->d_sine, plan->d_cosine, plan->d_cfilt,
                               plan->n_lblocks, s, 1, plan->d_raw);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_tpal_decode_separated: launch", e));
    }
    call_end(plan, s);
    return CHD_OK;
}

// ---------------------------------------------------------------------------
// Fixed batches: all jobs resident, one tile launch, one demod launch, graph.

}  // extern "C"

struct chd_tpal_batch {
    chd_tpal_plan *plan = nullptr;
    int n_units = 0, n_frames = 0;
    tpal::SlabJob *d_slab = nullptr;
    tpal::FieldJob2D *d_field = nullptr;
    tpal::FrameJob *d_frame = nullptr;
    float *d_partials = nullptr;
    float *d_raw = nullptr;        // uncored Y/U/V of every frame, with luma coring on
    unsigned *d_wq = nullptr;      // work counters of the persistent tile launches
    void *d_out = nullptr;
    bool border = false;
    int strip = tpal::kStrip3, groups = 0;   // strip: set at creation
    bool half = false;             // every slab alternates in line parity
    double tile_ms = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *d_spec = nullptr;        // learned mask: spectra and masks of chunk_units slabs
    float *d_mask = nullptr;
    int chunk_units = 0;
};

namespace {

constexpr int kBatchLearnedChunk = 16;   // slabs per spectra/network/apply round

void batch_release(chd_tpal_batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->cap) hipStreamDestroy(b->cap);
    void *ptrs[] = {b->d_slab, b->d_field, b->d_frame, b->d_partials, b->d_spec, b->d_mask, b->d_raw, b->d_wq};
    for (void *q : ptrs)
        if (q) hipFree(q);
}

void batch_tiles(chd_tpal_batch *b, hipStream_t s, int strip, int groups)
{
    chd_tpal_plan *plan = b->plan;
    const Geom &g = plan->g;
    if (plan->kind == CHD_DEC_PAL_2D) return;   // the demodulator filters the composite itself
    if (plan->kind != CHD_DEC_TRANSFORM_3D) {
        tpal::launch_tpal2d(g, b->d_field, b->n_units, plan->n_ty2_max, plan->d_thr32, plan->d_thr64, s);
        return;
    }
    if (plan->net == nullptr) {
        tpal::launch_tpal3d(g, b->d_slab, b->n_units, plan->d_thr32, plan->d_thr64, s, tpal::kTilePlain, nullptr,
                            nullptr, strip, groups, b->d_wq, b->half);
        return;
    }
    chd_mask_net *net = plan->net;
    const int64_t tps = (int64_t)g.n_ty3 * g.n_tx3;
    for (int u0 = 0; u0 < b->n_units; u0 += b->chunk_units) {
        const int cnt = std::min(b->chunk_units, b->n_units - u0);
        tpal::launch_tpal3d(g, b->d_slab + u0, cnt, plan->d_thr32, plan->d_thr64, s,
                            b->half ? tpal::kTileFeatures : tpal::kTileSpectra, b->d_spec, nullptr, strip, groups,
                            b->d_wq, b->half);
        mnet::launch_infer(b->d_spec, cnt * tps, net->theta, net->f, b->d_mask, net->cus, s, b->half);
        tpal::launch_tpal3d(g, b->d_slab + u0, cnt, plan->d_thr32, plan->d_thr64, s,
                            b->half ? tpal::kTileApplyFwd : tpal::kTileApply, b->d_spec, b->d_mask, strip, groups,
                            b->d_wq, b->half);
    }
}

void batch_all(chd_tpal_batch *b, hipStream_t s, int strip, int groups)
{
    chd_tpal_plan *plan = b->plan;
    if (b->border) tpal::launch_fill_border(plan->g, b->d_out, b->n_frames, s);
    batch_tiles(b, s, strip, groups);
    // d_raw is allocated whenever coring is on, so this launch cannot refuse
    (void)tpal::launch_demod(plan->g, b->d_frame, b->n_frames, plan->d_sine, plan->d_cosine, plan->d_cfilt,
                             plan->n_lblocks, s, plan_nchroma(plan), b->d_raw);
}

hipError_t batch_capture(chd_tpal_batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    hipError_t e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        batch_all(b, b->cap, b->strip, b->groups);
        const hipError_t le = hipGetLastError();
        e = hipStreamEndCapture(b->cap, &b->graph);
        if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    return e;
}

}  // namespace

extern "C" {

chd_status_t chd_tpal_batch_create(chd_tpal_plan_t *plan, const chd_tpal_sequence_t *seqs, int32_t n_seqs,
                                   void *d_out, int32_t autotune, chd_tpal_batch_t **out)
{
    static const char *W = "chd_tpal_batch_create";
    if (plan == nullptr || out == nullptr || n_seqs <= 0 || seqs == nullptr || d_out == nullptr)
        return chd::fail(CHD_E_INVALID_ARG, std::string(W) + ": null argument or no sequences");
    for (int i = 0; i < n_seqs; ++i) {
        const chd_tpal_sequence_t &q = seqs[i];
        if (q.count < 0 || q.first_frame < 0 || q.first_frame + q.count > q.n_frames ||
            (q.count > 0 && q.d_fields == nullptr))
            return chd::fail(CHD_E_OUT_OF_RANGE, std::string(W) + ": frame run outside its sequence");
    }
    hipSetDevice(plan->device);
    wait_staging(plan);
    const Geom &g = plan->g;
    const bool is3d = plan->kind == CHD_DEC_TRANSFORM_3D;
    const int64_t fb = frame_bytes(plan);
    std::vector<tpal::SlabJob> slabs;
    std::vector<tpal::FieldJob2D> fjobs;
    std::vector<tpal::FrameJob> frames;
    // partial offsets are assigned first; device pointers are added below
    for (int si = 0; si < n_seqs; ++si) {
        const chd_tpal_sequence_t &q = seqs[si];
        if (q.count == 0) continue;
        auto field_ptr = [&](int64_t fi) -> const uint16_t * {
            return (fi >= 0 && fi < 2 * q.n_frames) ? q.d_fields + fi * g.field_elems : nullptr;
        };
        const size_t unit0 = is3d ? slabs.size() : fjobs.size();
        if (is3d) {
            for (int64_t k = 0; k < q.count + 2; ++k) {
                tpal::SlabJob sj{};
                const int64_t slab = q.first_frame - 2 + k;
                for (int z = 0; z < 8; ++z) {
                    sj.field[z] = field_ptr(2 * slab + z);
                    sj.offset[z] = z & 1;
                }
                sj.partial = reinterpret_cast<float *>((unit0 + k) * plan->unit_floats);  // offset for now
                slabs.push_back(sj);
            }
        } else if (plan->kind == CHD_DEC_TRANSFORM_2D) {
            for (int64_t k = 0; k < 2 * q.count; ++k) {
                tpal::FieldJob2D jj{};
                jj.field = field_ptr(2 * q.first_frame + k);
                jj.offset = (int32_t)(k & 1);
                jj.partial = reinterpret_cast<float *>((unit0 + k) * plan->unit_floats);
                fjobs.push_back(jj);
            }
        }
        for (int64_t i = 0; i < q.count; ++i) {
            tpal::FrameJob fj{};
            fj.out = reinterpret_cast<void *>((int64_t)frames.size() * fb);  // offset for now
            for (int fsel = 0; fsel < 2; ++fsel) {
                fj.field[fsel] = field_ptr(2 * (q.first_frame + i) + fsel);
                fj.offset[fsel] = fsel;
                if (is3d) {
                    const size_t a = (unit0 + i) * plan->unit_floats + (2 + fsel) * 2 * g.part_plane;
                    const size_t bb = (unit0 + i + 2) * plan->unit_floats + fsel * 2 * g.part_plane;
                    fj.chroma[fsel][0] = reinterpret_cast<float *>(a);
                    fj.chroma[fsel][1] = reinterpret_cast<float *>(a + g.part_plane);
                    fj.chroma[fsel][2] = reinterpret_cast<float *>(bb);
                    fj.chroma[fsel][3] = reinterpret_cast<float *>(bb + g.part_plane);
                } else if (plan->kind == CHD_DEC_TRANSFORM_2D) {
                    const size_t a = (unit0 + 2 * i + fsel) * plan->unit_floats;
                    fj.chroma[fsel][0] = reinterpret_cast<float *>(a);
                    fj.chroma[fsel][1] = reinterpret_cast<float *>(a + g.part_plane);
                }
            }
            fj.nchroma = is3d ? 4 : (plan->kind == CHD_DEC_TRANSFORM_2D ? 2 : 0);
            frames.push_back(fj);
        }
    }
    if (frames.empty()) return chd::fail(CHD_E_INVALID_ARG, std::string(W) + ": no frames to decode");
    auto *b = new chd_tpal_batch();
    b->plan = plan;
    b->n_units = (int)(is3d ? slabs.size() : fjobs.size());
    b->n_frames = (int)frames.size();
    b->d_out = d_out;
    b->border = (g.out_w != g.aw || g.out_h != g.last - g.first + 1);
    if (is3d) b->strip = tpal::choose_strip3(g, b->n_units);
    hipError_t e = hipSuccess;
    if (b->n_units > 0)
        e = hipMalloc((void **)&b->d_partials, (size_t)b->n_units * plan->unit_floats * sizeof(float));
    auto rebase = [&](const float *off) { return b->d_partials + reinterpret_cast<size_t>(off); };
    if (e == hipSuccess) {
        for (auto &sj : slabs) sj.partial = rebase(sj.partial);
        for (auto &jj : fjobs) jj.partial = rebase(jj.partial);
        for (auto &fj : frames) {
            fj.out = (char *)d_out + reinterpret_cast<int64_t>(fj.out);
            for (int fsel = 0; fsel < 2; ++fsel)
                for (int k = 0; k < fj.nchroma; ++k) fj.chroma[fsel][k] = rebase(fj.chroma[fsel][k]);
        }
    }
    if (e == hipSuccess && g.ynr > 0.0f)
        e = hipMalloc((void **)&b->d_raw, (size_t)b->n_frames * tpal::raw_frame_floats_host(g) * sizeof(float));
    if (is3d) b->half = tpal::slabs_alternate(slabs.data(), (int)slabs.size());
    if (e == hipSuccess && is3d) e = hipMalloc((void **)&b->d_slab, slabs.size() * sizeof(tpal::SlabJob));
    if (e == hipSuccess && !fjobs.empty())
        e = hipMalloc((void **)&b->d_field, fjobs.size() * sizeof(tpal::FieldJob2D));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_frame, frames.size() * sizeof(tpal::FrameJob));
    if (e == hipSuccess && is3d)
        e = hipMemcpy(b->d_slab, slabs.data(), slabs.size() * sizeof(tpal::SlabJob), hipMemcpyHostToDevice);
    if (e == hipSuccess && !fjobs.empty())
        e = hipMemcpy(b->d_field, fjobs.data(), fjobs.size() * sizeof(tpal::FieldJob2D), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_frame, frames.data(), frames.size() * sizeof(tpal::FrameJob), hipMemcpyHostToDevice);
    if (e == hipSuccess && is3d && plan->net != nullptr) {
        b->chunk_units = std::min(kBatchLearnedChunk, b->n_units);
        const size_t tiles = (size_t)b->chunk_units * g.n_ty3 * g.n_tx3;
        e = hipMalloc(&b->d_spec, tiles * tpal::kTileBins * 2 * sizeof(float));
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_mask, tiles * mnet::kP * sizeof(float));
    }
    if (e == hipSuccess && is3d) e = hipMalloc((void **)&b->d_wq, tpal::kWorkCounters * sizeof(unsigned));
    if (e == hipSuccess && is3d) e = hipMemset(b->d_wq, 0, tpal::kWorkCounters * sizeof(unsigned));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    if (e != hipSuccess) {
        batch_release(b);
        delete b;
        return chd::fail(CHD_E_OOM, hip_msg(W, e));
    }

    // autotune: strip length x grid (one workgroup per item, or persistent)
    if (autotune && is3d) {
        int cus = 0;
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, plan->device);
        if (cus <= 0) cus = 256;
        const int strips[5] = {4, 8, 12, 16, 24};
        // 0: one workgroup per work item; otherwise persistent workgroups per
        // CU (12 fill the CU at the tile kernel's 3 waves per SIMD)
        const int per_cu[5] = {0, 4, 6, 8, 12};
        hipEvent_t t0 = nullptr, t1 = nullptr;
        e = hipEventCreate(&t0);
        if (e == hipSuccess) e = hipEventCreate(&t1);
        double best = 1e30;
        for (int si = 0; si < 5 && e == hipSuccess; ++si) {
            for (int pi = 0; pi < 5 && e == hipSuccess; ++pi) {
                const int groups = per_cu[pi] * cus;
                batch_tiles(b, b->cap, strips[si], groups);   // warm-up
                double t = 1e30;
                for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
                    hipEventRecord(t0, b->cap);
                    batch_tiles(b, b->cap, strips[si], groups);
                    hipEventRecord(t1, b->cap);
                    e = hipEventSynchronize(t1);
                    float ms = 0.0f;
                    hipEventElapsedTime(&ms, t0, t1);
                    t = std::min(t, (double)ms);
                }
                if (e == hipSuccess && t < best) {
                    best = t;
                    b->strip = strips[si];
                    b->groups = groups;
                }
            }
        }
        if (t0) hipEventDestroy(t0);
        if (t1) hipEventDestroy(t1);
        b->tile_ms = best < 1e30 ? best : 0.0;
        if (e != hipSuccess) {
            batch_release(b);
            delete b;
            return chd::fail(CHD_E_INTERNAL, hip_msg(W, e));
        }
    }

    // capture the whole batch once
    e = batch_capture(b);
    if (e != hipSuccess) {
        batch_release(b);
        delete b;
        return chd::fail(CHD_E_INTERNAL, hip_msg(W, e));
    }
    *out = b;
    return CHD_OK;
}

chd_status_t chd_tpal_batch_run(chd_tpal_batch_t *batch, void *stream)
{
    if (batch == nullptr) return chd::fail(CHD_E_INVALID_ARG, "chd_tpal_batch_run: null batch");
    hipSetDevice(batch->plan->device);
    const hipError_t e = hipGraphLaunch(batch->exec, (hipStream_t)stream);
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_tpal_batch_run", e));
    return CHD_OK;
}

chd_status_t chd_tpal_batch_config(const chd_tpal_batch_t *batch, int32_t *strip, int32_t *groups, double *tile_ms)
{
    if (batch == nullptr) return chd::fail(CHD_E_INVALID_ARG, "chd_tpal_batch_config: null batch");
    if (strip) *strip = batch->strip;
    if (groups) *groups = batch->groups;
    if (tile_ms) *tile_ms = batch->tile_ms;
    return CHD_OK;
}

chd_status_t chd_tpal_batch_set_shape(chd_tpal_batch_t *batch, int32_t strip, int32_t groups)
{
    if (batch == nullptr || strip < 4 || strip > 64 || strip % 4 != 0 || groups < 0)
        return chd::fail(CHD_E_INVALID_ARG,
                         "chd_tpal_batch_set_shape: strip must be a multiple of 4 from 4 to 64 and groups >= 0");
    hipSetDevice(batch->plan->device);
    hipStreamSynchronize(batch->cap);
    batch->strip = strip;
    batch->groups = groups;
    batch->tile_ms = 0.0;
    const hipError_t e = batch_capture(batch);
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_tpal_batch_set_shape", e));
    return CHD_OK;
}

void chd_tpal_batch_free(chd_tpal_batch_t *batch)
{
    if (batch == nullptr) return;
    hipSetDevice(batch->plan->device);
    if (batch->cap) hipStreamSynchronize(batch->cap);
    hipDeviceSynchronize();
    batch_release(batch);
    delete batch;
}

chd_status_t chd_tpal_set_timing(chd_tpal_plan_t *plan, int enable)
{
    if (plan == nullptr) return chd::fail(CHD_E_INVALID_ARG, "chd_tpal_set_timing: null argument");
    plan->timing = enable == 2 ? 2 : (enable != 0 ? 1 : 0);
    timing_reset(plan);
    return CHD_OK;
}

chd_status_t chd_tpal_last_timing(chd_tpal_plan_t *plan, double *tile_ms, double *demod_ms, int32_t *n_tile,
                                  int32_t *n_demod)
{
    if (plan == nullptr) return chd::fail(CHD_E_INVALID_ARG, "chd_tpal_last_timing: null argument");
    hipSetDevice(plan->device);
    timing_resolve(plan);
    if (tile_ms) *tile_ms = plan->last_tile_ms;
    if (demod_ms) *demod_ms = plan->last_demod_ms;
    if (n_tile) *n_tile = plan->n_tile_launches;
    if (n_demod) *n_demod = plan->n_demod_launches;
    return CHD_OK;
}

}  // extern "C"
