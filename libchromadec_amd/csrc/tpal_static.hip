// chd_tpal_decode without device allocation: a call's overlap-add partials,
// job lists and (with luma coring) uncored planes are carved out of one
// file-scope device array, `scratch`, instead of the plan's workspace.
//
// Each GPU holds its own copy of the array for the life of the process.
// Host code never passes the array's name to a kernel or a copy (on the host
// it is only a stand-in for the symbol); it uses the address that
// hipGetSymbolAddress returns for the current device, resolved once per
// device on the first call there.  Every call on a device shares that
// device's array, so calls that overlap on one GPU (any plan, any stream)
// race.  Nothing in the array needs to start at zero: the tile kernels
// write every partial sample the demodulator reads, and the demodulator
// writes every uncored sample the coring pass reads.
#include <hip/hip_runtime.h>

#include <chromadec/decoder.h>
#include <chromadec/tpal_device.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "tpal_device_internal.h"

constexpr size_t kScratchBytes = size_t(CHD_TPAL_STATIC_SCRATCH_MIB) << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<unsigned char *> g_scratch_ptr;   // per device ordinal: the array's device address

chd_status_t scratch_ptr(int device, unsigned char **out, const char *who)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device] == nullptr) {
        hipError_t e = hipSetDevice(device);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return chd::fail(CHD_E_INTERNAL, std::string(who) + ": hipGetSymbolAddress(scratch): " +
                                                 hipGetErrorName(e));
        g_scratch_ptr[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[device];
    return CHD_OK;
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Bytes of scratch a chunk of `frames` frames needs, and the workspace laid
// out from `base` (null base: sizes only).
size_t layout(const chd_tpal_plan *plan, int frames, unsigned char *base, TpalWorkspace *ws)
{
    const bool is3d = plan->kind == CHD_DEC_TRANSFORM_3D;
    const bool is2d = plan->kind == CHD_DEC_TRANSFORM_2D;
    // one run of F frames uses F + 2 slabs (3D) or 2F fields (2D); PAL 2D
    // keeps the per-field unit budget of the chunk loop but stores nothing
    const int units = is3d ? frames + 2 : 2 * frames;
    size_t off = 0;
    const size_t frame_off = off;
    off += align256((size_t)frames * sizeof(tpal::FrameJob));
    const size_t job_off = off;
    if (is3d) off += align256((size_t)units * sizeof(tpal::SlabJob));
    if (is2d) off += align256((size_t)units * sizeof(tpal::FieldJob2D));
    const size_t part_off = off;
    if (is3d || is2d) off += align256((size_t)units * plan->unit_floats * sizeof(float));
    const size_t raw_off = off;
    if (plan->g.ynr > 0.0f) off += align256((size_t)frames * tpal::raw_frame_floats_host(plan->g) * sizeof(float));
    if (base != nullptr && ws != nullptr) {
        ws->max_frames = frames;
        ws->max_units = units;
        ws->frame_jobs = reinterpret_cast<tpal::FrameJob *>(base + frame_off);
        ws->slab_jobs = is3d ? reinterpret_cast<tpal::SlabJob *>(base + job_off) : nullptr;
        ws->field_jobs = is2d ? reinterpret_cast<tpal::FieldJob2D *>(base + job_off) : nullptr;
        ws->partials = (is3d || is2d) ? reinterpret_cast<float *>(base + part_off) : nullptr;
        ws->raw = plan->g.ynr > 0.0f ? reinterpret_cast<float *>(base + raw_off) : nullptr;
    }
    return off;
}

}  // namespace

extern "C" chd_status_t chd_tpal_decode_static(chd_tpal_plan_t *plan, const chd_tpal_sequence_t *seqs,
                                               int32_t n_seqs, void *d_out, void *stream)
{
    static const char *W = "chd_tpal_decode_static";
    chd_status_t st = tpal_check_runs(plan, seqs, n_seqs, d_out, W);
    if (st != CHD_OK) return st;
    if (plan->net != nullptr)
        return chd::fail(CHD_E_UNSUPPORTED, std::string(W) +
                                                ": a plan with a learned mask needs tile spectra beyond the static "
                                                "scratch; use chd_tpal_decode");
    // the largest chunk that fits; the plan's pinned job staging bounds it too
    const size_t one = layout(plan, 1, nullptr, nullptr);
    if (one > kScratchBytes)
        return chd::fail(CHD_E_INVALID_ARG, std::string(W) + ": one frame of this geometry needs " +
                                                std::to_string(one) + " bytes of scratch, the static scratch holds " +
                                                std::to_string(kScratchBytes));
    int frames = 1;
    while (frames < plan->max_frames && layout(plan, frames + 1, nullptr, nullptr) <= kScratchBytes) ++frames;
    if (plan->kind == CHD_DEC_TRANSFORM_3D) frames = std::min(frames, plan->max_units - 2);
    unsigned char *base = nullptr;
    st = scratch_ptr(plan->device, &base, W);
    if (st != CHD_OK) return st;
    TpalWorkspace ws;
    layout(plan, frames, base, &ws);
    const hipError_t e = hipSetDevice(plan->device);
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, std::string(W) + ": hipSetDevice: " + hipGetErrorName(e));
    return tpal_decode_runs(plan, ws, seqs, n_seqs, d_out, (hipStream_t)stream, W);
}
