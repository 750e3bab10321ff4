// Internal view of a Transform-PAL device plan, shared by the device API
// (tpal_device.cpp) and the ABI decoder (decoder.cpp).
#pragma once

#include <chromadec/tpal_device.h>

#include <hip/hip_runtime_api.h>

#include <vector>

#include "tpal_internal.h"

struct chd_tpal_plan {
    tpal::Geom g{};
    int kind = 0;
    int device = 0;
    int n_ty2_max = 0;
    int n_lblocks = 0;
    float *d_thr32 = nullptr;
    double *d_thr64 = nullptr;
    float *d_sine = nullptr, *d_cosine = nullptr, *d_cfilt = nullptr;
    int max_frames = 0;
    int max_units = 0;          // slabs (3D) or fields (2D) per chunk
    size_t unit_floats = 0;
    float *d_partials = nullptr;
    tpal::SlabJob *d_slab_jobs = nullptr;
    tpal::FieldJob2D *d_field_jobs = nullptr;
    tpal::FrameJob *d_frame_jobs = nullptr;
    float *d_raw = nullptr;     // uncored Y/U/V of a chunk, with luma coring on
    tpal::SlabJob *h_slab_jobs = nullptr;       // pinned staging
    tpal::FieldJob2D *h_field_jobs = nullptr;
    tpal::FrameJob *h_frame_jobs = nullptr;
    hipEvent_t staging_done = nullptr;
    bool staging_pending = false;
    // The job lists of the last upload into d_slab_jobs / d_field_jobs and
    // d_frame_jobs (unit jobs, then frame jobs): a call that decodes the same
    // frames into the same buffers again skips the upload.
    std::vector<unsigned char> up_shadow;
    bool up_valid = false;
    // The partials, job lists and uncored planes are shared by every call on
    // the plan: a call on another stream first waits for the previous call's
    // last kernel (recorded in call_done on last_stream).
    hipEvent_t call_done = nullptr;
    hipStream_t last_stream = nullptr;
    bool call_pending = false;
    // Kernel timing: 0 off; 1 totals of the last decode call; 2 running
    // totals since chd_tpal_set_timing.  Each chunk records three events
    // (before the tile launch, between tile and demod launches, after the
    // demod) without waiting; they are summed when the totals are read.
    int timing = 0;
    std::vector<hipEvent_t> t_events;  // 3 per recorded chunk, created on first use
    int t_pending = 0;                 // chunks recorded and not yet summed
    double last_tile_ms = 0, last_demod_ms = 0;
    int n_tile_launches = 0, n_demod_launches = 0;
    // learned mask (3D): the network, and tile spectra / masks of one chunk
    struct chd_mask_net *net = nullptr;
    void *d_spec = nullptr;
    float *d_mask = nullptr;
};

// One chd_decode_frame window: the packed fields the reference's
// SourceField::loadFields would hand to PalDecoder::decodeFrames
// (14 for 3D with the frame at 4,5; 2 for 2D), nullptr for black padding.
struct TpalWindow {
    const uint16_t *field[14];
    int32_t offset[14];
    void *out;
};

// Device memory one decode call works in: overlap-add partials, job lists
// and, with luma coring, the uncored planes.  A plan owns one sized for
// max_batch_frames; chd_tpal_decode_static carves one out of a static
// device array.  max_units and max_frames never exceed the plan's (its
// pinned job staging is sized for those).
struct TpalWorkspace {
    float *partials = nullptr;          // max_units * plan->unit_floats
    int max_units = 0;                  // slabs (3D) or fields (2D) per chunk
    int max_frames = 0;                 // frames per chunk
    tpal::SlabJob *slab_jobs = nullptr;
    tpal::FieldJob2D *field_jobs = nullptr;
    tpal::FrameJob *frame_jobs = nullptr;
    float *raw = nullptr;               // max_frames raw planes, with luma coring on
};

chd_status_t tpal_plan_create_impl(const chd_tpal_params_t *p, chd_tpal_plan **out, const char *who);
chd_status_t tpal_decode_windows(chd_tpal_plan *plan, const TpalWindow *w, int n, hipStream_t s);
// chd_tpal_decode over an explicit workspace (arguments already checked).
chd_status_t tpal_decode_runs(chd_tpal_plan *plan, const TpalWorkspace &ws, const chd_tpal_sequence_t *seqs,
                              int n_seqs, void *d_out, hipStream_t s, const char *who);
// Validates the frame runs of a chd_tpal_decode-style call.
chd_status_t tpal_check_runs(const chd_tpal_plan *plan, const chd_tpal_sequence_t *seqs, int n_seqs, void *d_out,
                             const char *who);
