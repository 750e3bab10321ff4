// chd_decoder lifecycle, chd_decode_frame / chd_decode_frames_async and
// chd_frame accessors.
//
// Reference behaviour followed:
//   src/decoders/registry.cpp:85-173,208-234   option applicability, PAL config
//   src/abi/c_api_decoder.cpp:96-143           crop overrides and validation
//   src/abi/c_api_decoder.cpp:183-332          per-frame window (decodeSourceFrame)
//   src/abi/c_api_decoder.cpp:602-1049         commit
//   src/abi/c_api_decoder.cpp:1072-1152        sync and async decode
//   src/decoders/source_field.cpp:45-171       window loading, black padding
//   src/output/output_writer.cpp:195-239       output geometry / padding
//   src/abi/c_api_frame.cpp:24-238             frame accessors
//
// The chroma decode itself runs on the GPU through the device layer
// (tpal_device.cpp); this file owns no arithmetic on samples.
#include <chromadec/decoder.h>
#include <chromadec/tpal_mask.h>
#include <dlfcn.h>
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <tuple>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "dropout_internal.h"
#include "handles.h"
#include "tpal_device_internal.h"

namespace {

using chd::fail;

enum class OptType { F64, I32, Bool, Str };

constexpr int kMaxGpuWorkers = 8;   // host workers feeding one GPU in async decode (workers 1..8)
constexpr int kBatch = 4;           // consecutive frames one submission decodes together
constexpr int kCacheSlots = 14 * kBatch;   // packed fields resident per worker (a window needs 14)
// fields one submission stages from the file: a batch of consecutive frames
// and a 3D window's two frames behind and four ahead
constexpr int kStageFields = 2 * (kBatch + 6);
static_assert(kStageFields <= tpal::kPackListMax, "staged fields are packed by one launch");

bool is_pal_kind(chd_decoder_kind_t k)
{
    return k == CHD_DEC_PAL_2D || k == CHD_DEC_TRANSFORM_2D || k == CHD_DEC_TRANSFORM_3D;
}
bool is_comb_kind(chd_decoder_kind_t k)
{
    return k == CHD_DEC_NTSC_1D || k == CHD_DEC_NTSC_2D || k == CHD_DEC_NTSC_3D || k == CHD_DEC_NTSC_3D_NO_ADAPT ||
           k == CHD_DEC_NN_TRANSFORM3D;
}
bool is_ldzeug_kind(chd_decoder_kind_t k)
{
    return k == CHD_DEC_LDZEUG_COLOR_CNN || k == CHD_DEC_LDZEUG_LUMA_SEP || k == CHD_DEC_LDZEUG_LUMA_SEP_FRAME;
}

// registry::optionApplies (registry.cpp:113-173), plus this build's
// transform_mask / transform_mask_weights for CHD_DEC_TRANSFORM_3D.
// Packaged learned-mask weights: data/tpal3d_mask.bin beside the library's
// lib/ directory.
std::string default_mask_weights()
{
    Dl_info info;
    if (dladdr((void *)&chd_decoder_commit, &info) == 0 || info.dli_fname == nullptr) return std::string();
    std::string lib = info.dli_fname;
    const size_t slash = lib.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
    return dir + "/../data/tpal3d_mask.bin";
}

// Header check of a weights file (the format chd_mask_net_load reads).
bool mask_weights_ok(const std::string &path)
{
    if (path.empty()) return false;
    std::ifstream in(path, std::ios::binary);
    char magic[8] = {};
    int32_t count = 0;
    if (!in.read(magic, 8) || std::memcmp(magic, "CHDMASK1", 8) != 0) return false;
    if (!in.read(reinterpret_cast<char *>(&count), 4) || count != CHD_MASK_NET_PARAMS) return false;
    in.seekg(0, std::ios::end);
    return (int64_t)in.tellg() == 12 + (int64_t)count * 4;
}

bool option_applies(chd_decoder_kind_t kind, const std::string &n, OptType t)
{
    if (t == OptType::I32 && (n == "padding_multiple" || n == "thread_count")) return true;
    if (t == OptType::Str && (n == "output_format" || n == "output_clamp" || n == "color_difference_precision" ||
                              n == "broadcast_scaling_precision"))
        return true;
    if (t == OptType::Bool && (n == "output_y4m_headers" || n == "reverse_field_order")) return true;
    if (t == OptType::I32 && (n == "first_active_frame_line" || n == "last_active_frame_line" ||
                              n == "first_active_sample" || n == "last_active_sample"))
        return true;
    const bool comb = is_comb_kind(kind) || kind == CHD_DEC_AUTO;
    const bool pal = is_pal_kind(kind) || kind == CHD_DEC_AUTO;
    const bool ldzeug = is_ldzeug_kind(kind);
    const bool nn3d = kind == CHD_DEC_NN_TRANSFORM3D;
    const bool secam = kind == CHD_DEC_SECAM;
    const bool hvd = kind == CHD_DEC_HVD_2D || kind == CHD_DEC_HVD_3D;
    const bool tpal = kind == CHD_DEC_TRANSFORM_2D || kind == CHD_DEC_TRANSFORM_3D;
    if (t == OptType::F64 && n == "chroma_gain") return comb || pal || ldzeug || secam || hvd;
    if (t == OptType::F64 && n == "chroma_phase_deg") return comb || pal || ldzeug || hvd;
    if (t == OptType::F64 && n == "luma_nr_level") return comb || pal || kind == CHD_DEC_MONO;
    if (t == OptType::F64 && n == "chroma_nr_level") return comb;
    if (t == OptType::Bool && n == "phase_compensation") return comb || ldzeug;
    if (t == OptType::F64 && n == "comb_adapt_threshold") return kind == CHD_DEC_NTSC_3D;
    if (t == OptType::F64 && n == "comb_chroma_weight") return kind == CHD_DEC_NTSC_3D;
    if (t == OptType::Bool && n == "comb_show_map") return kind == CHD_DEC_NTSC_3D;
    if (t == OptType::Str && (n == "chroma_ident_mode" || n == "chroma_ident_manual")) return secam;
    if (t == OptType::F64 && (n == "chroma_click_nr_level" || n == "chroma_click_env_dip_db" ||
                              n == "chroma_click_freq_overshoot"))
        return secam;
    if (t == OptType::F64 && n == "transform_threshold") return tpal;
    if (t == OptType::Str && n == "transform_thresholds_file") return tpal;
    if (t == OptType::Str && (n == "transform_mask" || n == "transform_mask_weights"))
        return kind == CHD_DEC_TRANSFORM_3D;
    if (t == OptType::I32 && n == "hvd_cg_iterations") return hvd;
    if (t == OptType::F64 && n == "hvd_temporal_strength") return kind == CHD_DEC_HVD_3D;
    if (t == OptType::F64 && n == "nn_input_magnitude_scale") return nn3d;
    if (t == OptType::Bool && n == "nn_chroma_bandpass")
        return kind == CHD_DEC_LDZEUG_LUMA_SEP || kind == CHD_DEC_LDZEUG_LUMA_SEP_FRAME;
    return false;
}

std::string hip_err(hipError_t e) { return std::string(hipGetErrorName(e)) + " (" + hipGetErrorString(e) + ")"; }

// One worker: its own plan, stream, device field cache and staging, the
// equivalent of the reference's per-worker decoder instance (handles.h:136-148).
struct Worker {
    std::mutex mu;
    chd_tpal_plan *plan = nullptr;
    chd_mask_net_t *net = nullptr;      // learned mask (transform_mask "learned")
    hipStream_t stream = nullptr;
    uint16_t *d_slots = nullptr;
    std::vector<int64_t> slot_key;      // field number * 2 + shifted, -1 = empty
    std::vector<uint64_t> slot_tick;
    uint64_t tick = 0;
    uint16_t *d_stage = nullptr;        // staged fields on the device, before packing into slots
    // Two submissions in flight: submission slot s stages fields in
    // h_stage[s] (pinned), decodes up to kBatch frames into d_out[s] on
    // `stream`, and those come down on copy_stream, so one submission's
    // download overlaps the next one's decode.  A slot is reused only after
    // its download is waited for, which also orders its staging buffer.
    uint16_t *h_stage[2] = {nullptr, nullptr};
    const uint16_t *h_stage_dev[2] = {nullptr, nullptr};   // their device addresses
    void *d_out[2] = {nullptr, nullptr};
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_dec[2] = {nullptr, nullptr}, ev_copy[2] = {nullptr, nullptr};
    bool copy_busy[2] = {false, false};
    int device = 0;
    uint64_t seen_revision = 0;         // dropout options the cached fields were corrected with

    ~Worker()
    {
        if (plan) {
            hipSetDevice(device);
            if (stream) (void)hipStreamSynchronize(stream);
            chd_tpal_plan_free(plan);
        }
        chd_mask_net_free(net);
        if (copy_stream) (void)hipStreamSynchronize(copy_stream);
        if (d_slots) (void)hipFree(d_slots);
        for (void *o : d_out)
            if (o) (void)hipFree(o);
        if (d_stage) (void)hipFree(d_stage);
        for (uint16_t *h : h_stage)
            if (h) (void)hipHostFree(h);
        for (hipEvent_t e : {ev_dec[0], ev_dec[1], ev_copy[0], ev_copy[1]})
            if (e) (void)hipEventDestroy(e);
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

}  // namespace

// Idle pinned blocks kept (freeing one synchronises the device): async
// decode holds up to two batch blocks per worker in flight.
constexpr size_t kPoolKeepOne = 32;
constexpr size_t kPoolKeepBatch = 2 * kMaxGpuWorkers + 2;

FramePool::~FramePool()
{
    for (void *p : free_one) (void)hipHostFree(p);
    for (void *p : free_batch) (void)hipHostFree(p);
}

std::shared_ptr<uint8_t> FramePool::get(int n)
{
    const bool one = n == 1;
    std::vector<void *> &list = one ? free_one : free_batch;
    void *p = nullptr;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!list.empty()) {
            p = list.back();
            list.pop_back();
        }
    }
    if (p == nullptr && hipHostMalloc(&p, bytes * (one ? 1 : batch), 0) != hipSuccess) return nullptr;
    auto self = shared_from_this();
    return std::shared_ptr<uint8_t>(static_cast<uint8_t *>(p), [self, one](uint8_t *q) {
        {
            std::lock_guard<std::mutex> lk(self->mu);
            std::vector<void *> &l = one ? self->free_one : self->free_batch;
            if (l.size() < (one ? kPoolKeepOne : kPoolKeepBatch)) {
                l.push_back(q);
                return;
            }
        }
        (void)hipHostFree(q);
    });
}

struct chd_decoder {
    chd_video_t *video = nullptr;
    chd_decoder_kind_t kind = CHD_DEC_AUTO;
    std::unordered_map<std::string, double> f64;
    std::unordered_map<std::string, int32_t> i32;
    std::unordered_map<std::string, bool> boolean;
    std::unordered_map<std::string, std::string> str;
    bool committed = false;
    std::shared_ptr<FramePool> framePool;   // pinned host frames (set at commit)

    // committed state
    chd_decoder_kind_t resolved = CHD_DEC_AUTO;
    bool reverseFieldOrder = false;
    chd::VideoParams vp;
    chd_pixel_format_t format = CHD_PIXEL_YUV444P16;
    int32_t outW = 0, outH = 0, padL = 0, padT = 0, planes = 3;
    int32_t lookBehind = 0, lookAhead = 0;
    int32_t threadCount = 1;
    chd_tpal_params_t params{};
    std::vector<double> thresholds;
    std::string maskWeights;            // non-empty: learned mask from this file

    // dropout concealment (dropout.h)
    std::mutex dropMu;
    chd_dropout_opts_t dropoutOpts{};
    bool dropoutSet = false;
    uint64_t dropoutRevision = 1;
    chd_dropout_stats_t lastStats{};
    chd::dropout::Params dropParams;
    struct Corrected {
        std::vector<uint16_t> first, second;
        chd::dropout::Stats stats;
    };
    // corrected frames by (first field, second field, options revision), so
    // the frames of overlapping windows are corrected once
    std::map<std::tuple<int32_t, int32_t, uint64_t>, std::shared_ptr<const Corrected>> corrected;
    std::deque<std::tuple<int32_t, int32_t, uint64_t>> correctedOrder;
    std::mutex workersMu;
    std::vector<std::unique_ptr<Worker>> workers;

    bool firstFieldFirst() const { return video->meta.isFirstFieldFirst != reverseFieldOrder; }
};

namespace {

template <typename T>
chd_status_t set_opt(chd_decoder_t *d, const char *name, OptType t, T value,
                     std::unordered_map<std::string, T> chd_decoder::*dst)
{
    if (d == nullptr || name == nullptr) return fail(CHD_E_INVALID_ARG, "chromadec: chd_decoder_set_option: null argument");
    if (d->committed) return fail(CHD_E_INVALID_ARG, "chd_decoder_set_option: decoder already committed");
    if (!option_applies(d->kind, name, t))
        return fail(CHD_E_INVALID_ARG,
                    std::string("chd_decoder_set_option: option \"") + name + "\" not valid for this decoder kind");
    (d->*dst)[name] = std::move(value);
    return CHD_OK;
}

chd_status_t worker_init(chd_decoder_t *d, Worker &w, const char *who)
{
    if (w.plan != nullptr) return CHD_OK;
    chd_tpal_params_t p = d->params;
    p.max_batch_frames = kBatch;
    chd_status_t st = tpal_plan_create_impl(&p, &w.plan, who);
    if (st != CHD_OK) return st;
    w.device = p.device;
    const tpal::Geom &g = w.plan->g;
    hipError_t e = hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&w.d_slots, (size_t)kCacheSlots * g.field_elems * sizeof(uint16_t));
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipMalloc(&w.d_out[k], kBatch * chd_tpal_frame_bytes(w.plan));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&w.copy_stream, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipEventCreateWithFlags(&w.ev_dec[k], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&w.ev_copy[k], hipEventDisableTiming);
    }
    const size_t stage_bytes = (size_t)kStageFields * g.fw * g.fh * sizeof(uint16_t);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipHostMalloc((void **)&w.h_stage[k], stage_bytes, 0);
        if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&w.h_stage_dev[k], w.h_stage[k], 0);
    }
    if (e == hipSuccess) e = hipMalloc((void **)&w.d_stage, stage_bytes);
    if (e == hipSuccess)
        e = hipMemsetD16Async((hipDeviceptr_t)w.d_slots, (unsigned short)d->vp.black16bIre,
                              (size_t)kCacheSlots * g.field_elems, w.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(w.stream);
    if (e != hipSuccess) return fail(CHD_E_OOM, std::string(who) + ": device allocation failed: " + hip_err(e));
    if (!d->maskWeights.empty()) {
        // inference only: no training working memory
        st = chd_mask_net_create(p.device, 1, &w.net);
        if (st == CHD_OK) st = chd_mask_net_load(w.net, d->maskWeights.c_str());
        if (st == CHD_OK) st = chd_tpal_set_mask_net(w.plan, w.net);
        if (st != CHD_OK) return st;
    }
    w.slot_key.assign(kCacheSlots, -1);
    w.slot_tick.assign(kCacheSlots, 0);
    return CHD_OK;
}

Worker *get_worker(chd_decoder_t *d, size_t idx)
{
    std::lock_guard<std::mutex> lk(d->workersMu);
    while (d->workers.size() <= idx) d->workers.push_back(std::make_unique<Worker>());
    return d->workers[idx].get();
}

// Fields a submission reads from the file into its pinned staging buffer;
// flush_stage uploads them in one copy and packs them into their cache slots
// in one launch.
struct Staging {
    uint16_t *h = nullptr;
    const uint16_t *h_dev = nullptr;   // h's device address
    int n = 0;
    tpal::PackList list{};
};

chd_status_t flush_stage(Worker &w, Staging &sg)
{
    if (sg.n == 0) return CHD_OK;
    const tpal::Geom &g = w.plan->g;
    const int64_t fs = (int64_t)g.fw * g.fh;
    // One contiguous copy, then the pack kernel lays each field out with its
    // margins (row-by-row pitched copies cost a DMA command per line).  A
    // kernel reads the pinned staging buffer: the DMA engine that carries
    // decoded frames down would otherwise carry the fields up in turn.
    tpal::launch_copy_in(sg.h_dev, w.d_stage, (size_t)sg.n * fs * sizeof(uint16_t), w.stream);
    sg.list.n = sg.n;
    tpal::launch_pack_list(g, w.d_stage, fs, w.d_slots, sg.list, w.stream);
    const hipError_t e = hipGetLastError();
    sg.n = 0;
    if (e != hipSuccess) return fail(CHD_E_INTERNAL, "chd_decode_frame: field upload failed: " + hip_err(e));
    return CHD_OK;
}

// Device copy of field `number` (1-based), read from the file, or, when
// `corrected` is given, that host copy (already shifted and concealed).  A
// concealed field depends on the frame it was concealed in, so its key
// holds the partner field: with both fields marked first, one field can be
// the second field of one frame and the first of the next.  `pinned` marks
// the slots the current submission uses; a new field is staged in `sg`.
const uint16_t *cache_field(chd_decoder_t *d, Worker &w, int32_t number, bool shifted, Staging &sg,
                            std::vector<int> &pinned, chd_status_t &st, const uint16_t *corrected = nullptr,
                            int32_t partner = 0)
{
    const tpal::Geom &g = w.plan->g;
    // bit 0: concealed copy; bit 1: shifted; the field number above, and for
    // a concealed copy the partner field number in the upper 32 bits
    const int64_t key = corrected ? (((int64_t)partner << 32) | ((int64_t)number << 2) | (shifted ? 2 : 0) | 1)
                                  : (((int64_t)number << 2) | (shifted ? 2 : 0));
    w.tick++;
    for (int s = 0; s < kCacheSlots; ++s) {
        if (w.slot_key[s] == key) {
            w.slot_tick[s] = w.tick;
            pinned.push_back(s);
            return w.d_slots + (size_t)s * g.field_elems;
        }
    }
    int victim = -1;
    for (int s = 0; s < kCacheSlots; ++s) {
        if (std::find(pinned.begin(), pinned.end(), s) != pinned.end()) continue;
        if (victim < 0 || w.slot_tick[s] < w.slot_tick[victim]) victim = s;
    }
    if (victim < 0) {
        st = fail(CHD_E_INTERNAL, "chd_decode_frame: field cache exhausted");
        return nullptr;
    }
    if (sg.n == kStageFields) {
        // staging full (only with unusual field numbering): upload what is
        // staged and wait for the copy before refilling the buffer
        st = flush_stage(w, sg);
        if (st == CHD_OK && hipStreamSynchronize(w.stream) != hipSuccess)
            st = fail(CHD_E_INTERNAL, "chd_decode_frame: field upload failed");
        if (st != CHD_OK) return nullptr;
    }
    uint16_t *stage = sg.h + (size_t)sg.n * g.fw * g.fh;
    if (corrected != nullptr) {
        std::memcpy(stage, corrected, (size_t)g.fw * g.fh * sizeof(uint16_t));
    } else if (!d->video->tbc.read_field(number, stage)) {
        st = fail(CHD_E_IO, "chd_decode_frame: failed to read field " + std::to_string(number));
        return nullptr;
    }
    // SourceField::shiftSecondField (source_field.cpp:51-66) drops the first
    // two samples of the field and appends two black ones: shift 2.
    sg.list.slot[sg.n] = victim;
    sg.list.shift[sg.n] = shifted && corrected == nullptr ? 2 : 0;
    sg.n++;
    w.slot_key[victim] = key;
    w.slot_tick[victim] = w.tick;
    pinned.push_back(victim);
    return w.d_slots + (size_t)victim * g.field_elems;
}

chd::dropout::Params dropout_params(const chd::VideoParams &vp)
{
    chd::dropout::Params p;
    p.field_width = vp.fieldWidth;
    p.field_height = vp.fieldHeight;
    p.colour_burst_end = vp.colourBurstEnd;
    p.active_video_end = vp.activeVideoEnd;
    p.first_active_frame_line = vp.firstActiveFrameLine;
    p.last_active_frame_line = vp.lastActiveFrameLine;
    p.system = (int32_t)vp.system;
    return p;
}

std::vector<chd::dropout::Run> runs_of(const chd::DropOuts &d)
{
    std::vector<chd::dropout::Run> r(d.startx.size());
    for (size_t i = 0; i < r.size(); ++i) r[i] = {d.startx[i], d.endx[i], d.fieldLine[i]};
    return r;
}

// A frame's fields with their dropouts concealed (c_api_decoder.cpp:
// 238-310): read, shift the second field of a subcarrier-locked PAL source
// with its dropouts (source_field.cpp:51-66), correct.
std::shared_ptr<const chd_decoder::Corrected> corrected_frame(chd_decoder_t *d, int32_t f1, int32_t f2, bool shift,
                                                              const chd_dropout_opts_t &opts, uint64_t rev,
                                                              const chd::dropout::Params &params, chd_status_t &st)
{
    const auto key = std::make_tuple(f1, f2, rev);
    {
        std::lock_guard<std::mutex> lk(d->dropMu);
        auto it = d->corrected.find(key);
        if (it != d->corrected.end()) return it->second;
    }
    const chd::VideoParams &vp = d->video->meta.vp;
    const size_t n = (size_t)vp.fieldWidth * vp.fieldHeight;
    auto c = std::make_shared<chd_decoder::Corrected>();
    c->first.resize(n);
    c->second.resize(n);
    if (!d->video->tbc.read_field(f1, c->first.data()) || !d->video->tbc.read_field(f2, c->second.data())) {
        st = fail(CHD_E_IO, "chd_decode_frame: failed to read fields " + std::to_string(f1) + "/" + std::to_string(f2));
        return nullptr;
    }
    std::vector<chd::dropout::Run> r1 = runs_of(d->video->meta.field(f1).dropOuts);
    std::vector<chd::dropout::Run> r2 = runs_of(d->video->meta.field(f2).dropOuts);
    if (shift) {
        std::memmove(c->second.data(), c->second.data() + 2, (n - 2) * sizeof(uint16_t));
        c->second[n - 2] = c->second[n - 1] = (uint16_t)vp.black16bIre;
        std::vector<chd::dropout::Run> s;
        for (const auto &r : r2)
            if (r.endx - 2 > 0) s.push_back({std::max(r.startx - 2, 0), r.endx - 2, r.field_line});
        r2.swap(s);
    }
    chd::dropout::correct_frame(params, c->first.data(), c->second.data(), r1, r2, opts.overcorrect != 0,
                                opts.intra_field_only != 0, &c->stats);
    std::lock_guard<std::mutex> lk(d->dropMu);
    d->corrected[key] = c;
    d->correctedOrder.push_back(key);
    while (d->correctedOrder.size() > 48) {
        d->corrected.erase(d->correctedOrder.front());
        d->correctedOrder.pop_front();
    }
    return c;
}

// Wait for `ev` by polling: a blocking wait sleeps until an interrupt, and
// with several workers each waiting on its own frames those wake-ups came
// milliseconds late, leaving the GPU idle between frames.  Polls for up to
// 50 ms, then blocks.
hipError_t wait_event(hipEvent_t ev)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t e = hipEventQuery(ev);
        if (e != hipErrorNotReady) return e;
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) return hipEventSynchronize(ev);
        std::this_thread::yield();
    }
}

// One submission: up to kBatch frames, each with its status and, once
// decoded, its pinned host frame.
struct Submission {
    int n = 0;
    int64_t idx[kBatch] = {};
    chd_status_t st[kBatch] = {};
    std::unique_ptr<chd_frame> frame[kBatch];
    bool busy = false;
};

// SourceField::loadFields (source_field.cpp:68-171) for one frame: the
// window's fields, made resident in the worker's cache (new ones staged in
// `sg`), and the field order offsets.
chd_status_t load_window(chd_decoder_t *d, Worker &w, int64_t frame_index, const chd_dropout_opts_t &dopts,
                         uint64_t drev, Staging &sg, std::vector<int> &pinned, TpalWindow &win)
{
    chd::Metadata &meta = d->video->meta;
    const bool fff = d->firstFieldFirst();
    const int32_t numFrames = meta.numberOfFrames(fff);
    if (frame_index < 0 || frame_index >= numFrames)
        return fail(CHD_E_OUT_OF_RANGE, "chd_decode_frame: frame_index out of range");
    const int32_t behind = d->lookBehind, ahead = d->lookAhead;
    const int32_t nwin = 2 * (behind + 1 + ahead);
    const int32_t firstFrameNumber = (int32_t)frame_index + 1;
    const int32_t numFields = meta.numberOfFields();
    const int32_t numSource = d->video->tbc.available_fields;
    const bool shiftSecond = (d->vp.system == chd::System::PAL || d->vp.system == chd::System::PAL_M) &&
                             d->vp.isSubcarrierLocked;
    chd_status_t st = CHD_OK;
    std::memset(&win, 0, sizeof(win));
    chd::dropout::Stats frameStats;
    int32_t frameNumber = firstFrameNumber - behind;
    for (int32_t i = 0; i < nwin; i += 2, ++frameNumber) {
        bool blank = frameNumber < 1 || frameNumber > numFrames;
        int32_t mframe = blank ? 1 : frameNumber;
        int32_t f1 = meta.fieldNumber(mframe, false, fff), f2 = meta.fieldNumber(mframe, true, fff);
        auto metaSafe = [&](int32_t a, int32_t b) { return a >= 1 && b >= 1 && a <= numFields && b <= numFields; };
        bool metaBad = !metaSafe(f1, f2);
        const bool srcBad = numSource != -1 && (f1 > numSource || f2 > numSource);
        if (!blank && (metaBad || srcBad)) {
            chd::log_message(CHD_LOG_WARN, "chd_decode_frame: frame " + std::to_string(frameNumber) +
                                               " is outside the readable fields; decoding it as black");
            blank = true;
            mframe = 1;
            f1 = meta.fieldNumber(mframe, false, fff);
            f2 = meta.fieldNumber(mframe, true, fff);
            metaBad = !metaSafe(f1, f2);
        }
        if (!metaBad) {
            win.offset[i] = meta.field(f1).isFirstField ? 0 : 1;
            win.offset[i + 1] = meta.field(f2).isFirstField ? 0 : 1;
        } else {
            win.offset[i] = 0;
            win.offset[i + 1] = 1;
            blank = true;
        }
        const bool hasDrops = !metaBad && (!meta.field(f1).dropOuts.empty() || !meta.field(f2).dropOuts.empty());
        if (blank) {
            win.field[i] = nullptr;
            win.field[i + 1] = nullptr;
        } else if (drev != 0 && hasDrops) {
            // correct every frame of the window (c_api_decoder.cpp:297-315)
            auto c = corrected_frame(d, f1, f2, shiftSecond, dopts, drev, d->dropParams, st);
            if (st != CHD_OK) return st;
            if (i == 2 * behind) frameStats = c->stats;
            win.field[i] = cache_field(d, w, f1, false, sg, pinned, st, c->first.data(), f2);
            if (st != CHD_OK) return st;
            win.field[i + 1] = cache_field(d, w, f2, shiftSecond, sg, pinned, st, c->second.data(), f1);
            if (st != CHD_OK) return st;
        } else {
            win.field[i] = cache_field(d, w, f1, false, sg, pinned, st);
            if (st != CHD_OK) return st;
            win.field[i + 1] = cache_field(d, w, f2, shiftSecond, sg, pinned, st);
            if (st != CHD_OK) return st;
        }
    }
    // The decoded frame's own fields are black (nullptr) when the metadata or
    // the .tbc do not cover it, as in SourceField::loadFields; the kernels
    // decode a nullptr field as black.
    std::lock_guard<std::mutex> lk(d->dropMu);
    d->lastStats.corrected = frameStats.corrected;
    d->lastStats.failed = frameStats.failed;
    d->lastStats.total_distance = frameStats.total_distance;
    return CHD_OK;
}

// decodeFrameLocked (c_api_decoder.cpp:183-332, 461-515) for frames
// idx[0..nb) on worker `w`, in two halves: submit_frames loads the windows,
// enqueues one decode of all of them into submission slot `slot` and the
// downloads into pinned frames; finish_frames waits for the downloads.  A
// worker holds one submission per slot; a frame whose window cannot be
// loaded keeps its error status and the others still decode.
void submit_frames(chd_decoder_t *d, Worker &w, const int64_t *idx, int nb, int slot, Submission &sub)
{
    sub.n = nb;
    sub.busy = true;
    for (int j = 0; j < nb; ++j) {
        sub.idx[j] = idx[j];
        sub.st[j] = CHD_OK;
        sub.frame[j].reset();
    }
    auto fail_all = [&](chd_status_t st) {
        for (int j = 0; j < nb; ++j)
            if (sub.st[j] == CHD_OK) sub.st[j] = st;
    };
    chd_status_t st = worker_init(d, w, "chd_decode_frame");
    if (st != CHD_OK) return fail_all(st);
    hipSetDevice(w.device);

    chd_dropout_opts_t dopts{};
    uint64_t drev = 0;
    {
        std::lock_guard<std::mutex> lk(d->dropMu);
        if (d->dropoutSet && d->dropoutOpts.enabled) {
            dopts = d->dropoutOpts;
            drev = d->dropoutRevision;
        }
    }
    if (drev != w.seen_revision) {
        // cached device copies of corrected fields belong to other options
        for (int s = 0; s < kCacheSlots; ++s)
            if (w.slot_key[s] >= 0 && (w.slot_key[s] & 1)) w.slot_key[s] = -1;
        w.seen_revision = drev;
    }
    const size_t fb = chd_tpal_frame_bytes(w.plan);
    Staging sg;
    sg.h = w.h_stage[slot];
    sg.h_dev = w.h_stage_dev[slot];
    std::vector<int> pinned;
    TpalWindow win[kBatch];
    int at[kBatch];   // frame of each decoded window
    int nw = 0;
    for (int j = 0; j < nb; ++j) {
        sub.st[j] = load_window(d, w, idx[j], dopts, drev, sg, pinned, win[nw]);
        if (sub.st[j] != CHD_OK) continue;
        win[nw].out = static_cast<char *>(w.d_out[slot]) + (size_t)nw * fb;
        at[nw++] = j;
    }
    st = flush_stage(w, sg);
    if (st == CHD_OK && nw > 0) {
        // the slot's previous frames must have come down before they are
        // overwritten
        if (w.copy_busy[slot]) (void)hipStreamWaitEvent(w.stream, w.ev_copy[slot], 0);
        st = tpal_decode_windows(w.plan, win, nw, w.stream);
    }
    if (st != CHD_OK) {
        // nothing may still read the staging buffer once the slot is reused
        (void)hipStreamSynchronize(w.stream);
        return fail_all(st);
    }
    if (nw == 0) {
        // no decode follows: wait for the staged fields to be read here, as
        // the slot's staging buffer may be refilled next
        (void)hipStreamSynchronize(w.stream);
        return;
    }

    const bool is_float = d->format == CHD_PIXEL_YUV444PS || d->format == CHD_PIXEL_GRAYS ||
                          d->format == CHD_PIXEL_RGBS;
    const size_t bytes = (size_t)d->planes * d->outW * d->outH * (is_float ? sizeof(float) : sizeof(uint16_t));
    hipError_t e = hipEventRecord(w.ev_dec[slot], w.stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(w.copy_stream, w.ev_dec[slot], 0);
    // a full batch comes down in one copy into one block (the decoded frames
    // lie back to back in d_out); otherwise each frame in its own block
    const bool whole = nw == kBatch && bytes == fb;
    std::shared_ptr<uint8_t> block;
    if (whole) {
        block = d->framePool->get(kBatch);
        if (block == nullptr) {
            for (int k = 0; k < nw; ++k) sub.st[at[k]] = fail(CHD_E_OOM, "chd_decode_frame: no pinned host memory for the frame");
            nw = 0;
        } else if (e == hipSuccess) {
            e = hipMemcpyAsync(block.get(), w.d_out[slot], (size_t)nw * fb, hipMemcpyDeviceToHost, w.copy_stream);
        }
    }
    for (int k = 0; k < nw && e == hipSuccess; ++k) {
        const int j = at[k];
        auto frame = std::make_unique<chd_frame>();
        frame->format = d->format;
        frame->width = d->outW;
        frame->height = d->outH;
        frame->planes = d->planes;
        frame->info.format = d->format;
        frame->info.width = d->outW;
        frame->info.height = d->outH;
        frame->info.num_planes = d->planes;
        frame->info.frame_index = idx[j];
        // straight into a pinned block: DMA at link rate
        if (whole) {
            frame->block = block;
            frame->data = block.get() + (size_t)k * fb;
        } else {
            frame->block = d->framePool->get(1);
            if (frame->block == nullptr) {
                sub.st[j] = fail(CHD_E_OOM, "chd_decode_frame: no pinned host memory for the frame");
                continue;
            }
            frame->data = frame->block.get();
            e = hipMemcpyAsync(frame->data, static_cast<char *>(w.d_out[slot]) + (size_t)k * fb, bytes,
                               hipMemcpyDeviceToHost, w.copy_stream);
        }
        sub.frame[j] = std::move(frame);
    }
    if (e == hipSuccess) e = hipEventRecord(w.ev_copy[slot], w.copy_stream);
    if (e != hipSuccess) {
        // the pinned buffers go back to the pool only once nothing writes them
        (void)hipStreamSynchronize(w.copy_stream);
        for (auto &f : sub.frame) f.reset();
        return fail_all(fail(CHD_E_INTERNAL, "chd_decode_frame: GPU decode failed: " + hip_err(e)));
    }
    w.copy_busy[slot] = true;
}

// Waits for submission slot `slot`'s downloads and hands its frames, in
// order, to take(status, frame_index, frame).
template <class Take>
void finish_frames(Worker &w, int slot, Submission &sub, Take &&take)
{
    hipError_t e = hipSuccess;
    if (w.copy_busy[slot]) {
        e = wait_event(w.ev_copy[slot]);
        w.copy_busy[slot] = false;
    }
    const chd_status_t gpu = e == hipSuccess ? CHD_OK
                                             : fail(CHD_E_INTERNAL, "chd_decode_frame: GPU decode failed: " + hip_err(e));
    for (int j = 0; j < sub.n; ++j) {
        chd_status_t st = sub.st[j] != CHD_OK ? sub.st[j] : gpu;
        chd_frame_t *f = nullptr;
        if (st == CHD_OK) f = sub.frame[j].release();
        sub.frame[j].reset();
        take(st, sub.idx[j], f);
    }
    sub.busy = false;
}

chd_status_t decode_locked(chd_decoder_t *d, Worker &w, int64_t frame_index, chd_frame_t **out)
{
    Submission sub;
    submit_frames(d, w, &frame_index, 1, 0, sub);
    chd_status_t rc = CHD_OK;
    finish_frames(w, 0, sub, [&](chd_status_t st, int64_t, chd_frame_t *f) {
        rc = st;
        *out = f;
    });
    return rc;
}

int parse_format(const std::string &n, chd_pixel_format_t &f)
{
    // 1 = supported, 0 = known but not produced by this build, -1 = unknown
    if (n == "yuv444p16" || n == "YUV444P16") { f = CHD_PIXEL_YUV444P16; return 1; }
    if (n == "yuv444ps" || n == "YUV444PS") { f = CHD_PIXEL_YUV444PS; return 1; }
    if (n == "gray16" || n == "GRAY16") { f = CHD_PIXEL_GRAY16; return 1; }
    if (n == "grays" || n == "GRAYS") { f = CHD_PIXEL_GRAYS; return 1; }
    if (n == "rgb48" || n == "RGB48") { f = CHD_PIXEL_RGB48; return 1; }
    if (n == "rgbs" || n == "RGBS") { f = CHD_PIXEL_RGBS; return 1; }
    if (n == "yuv440p16" || n == "YUV440P16") { f = CHD_PIXEL_YUV440P16; return 0; }
    if (n == "yuv440ps" || n == "YUV440PS") { f = CHD_PIXEL_YUV440PS; return 0; }
    return -1;
}

int parse_clamp(const std::string &n)
{
    if (n == "none" || n == "NONE") return CHD_CLAMP_NONE;
    if (n == "legal_rgb_sdr" || n == "LEGAL_RGB_SDR") return CHD_CLAMP_LEGAL_RGB_SDR;
    if (n == "legal_rgb_hdr" || n == "LEGAL_RGB_HDR") return CHD_CLAMP_LEGAL_RGB_HDR;
    if (n == "legal_ycbcr_bt601" || n == "LEGAL_YCBCR_BT601") return CHD_CLAMP_LEGAL_YCBCR_BT601;
    return -1;
}

// color_conversion.cpp:16-67
bool color_scales(const std::string &cdp, const std::string &bsp, double &ecb, double &ecr, double &kr, double &kb)
{
    double ur, vr;
    if (cdp.empty() || cdp == "modern") { kr = 0.299; kb = 0.114; }
    else if (cdp == "classic") { kr = 0.30; kb = 0.11; }
    else return false;
    if (bsp.empty() || bsp == "scientific") { ur = 0.49211104112248356308804691718185; vr = 0.87728321993817866838972487283129; }
    else if (bsp == "modern") { ur = 0.492111; vr = 0.877283; }
    else if (bsp == "classic") { ur = 0.493; vr = 0.877; }
    else return false;
    ecb = 1.0 / (2.0 * (1.0 - kb) * ur);
    ecr = 1.0 / (2.0 * (1.0 - kr) * vr);
    return true;
}

}  // namespace

extern "C" {

chd_status_t chd_decoder_create(chd_video_t *v, chd_decoder_kind_t kind, chd_decoder_t **out)
{
    if (v == nullptr || out == nullptr) return fail(CHD_E_INVALID_ARG, "chromadec: chd_decoder_create: null argument");
    auto d = std::make_unique<chd_decoder>();
    d->video = v;
    d->kind = kind;
    *out = d.release();
    return CHD_OK;
}

void chd_decoder_free(chd_decoder_t *d) { delete d; }

chd_status_t chd_decoder_set_option_f64(chd_decoder_t *d, const char *name, double v)
{
    return set_opt(d, name, OptType::F64, v, &chd_decoder::f64);
}
chd_status_t chd_decoder_set_option_i32(chd_decoder_t *d, const char *name, int32_t v)
{
    return set_opt(d, name, OptType::I32, v, &chd_decoder::i32);
}
chd_status_t chd_decoder_set_option_bool(chd_decoder_t *d, const char *name, int v)
{
    return set_opt<bool>(d, name, OptType::Bool, v != 0, &chd_decoder::boolean);
}
chd_status_t chd_decoder_set_option_str(chd_decoder_t *d, const char *name, const char *v)
{
    if (v == nullptr) return fail(CHD_E_INVALID_ARG, "chromadec: chd_decoder_set_option_str: null argument (value)");
    return set_opt<std::string>(d, name, OptType::Str, v, &chd_decoder::str);
}

chd_status_t chd_decoder_has_option(const chd_decoder_t *d, const char *name)
{
    if (d == nullptr || name == nullptr) return fail(CHD_E_INVALID_ARG, "chromadec: chd_decoder_has_option: null argument");
    const bool any = option_applies(d->kind, name, OptType::F64) || option_applies(d->kind, name, OptType::I32) ||
                     option_applies(d->kind, name, OptType::Bool) || option_applies(d->kind, name, OptType::Str);
    return any ? CHD_OK : CHD_E_INVALID_ARG;
}

chd_status_t chd_decoder_commit(chd_decoder_t *d)
{
    const std::string fn = "chd_decoder_commit";
    if (d == nullptr) return fail(CHD_E_INVALID_ARG, "chromadec: chd_decoder_commit: null argument");
    if (d->committed) return CHD_OK;
    if (d->video == nullptr) return fail(CHD_E_INVALID_ARG, fn + ": video has no source");
    const chd::Metadata &meta = d->video->meta;
    const chd::System system = meta.vp.system;

    // registry::resolveAuto (registry.cpp:96-107)
    chd_decoder_kind_t resolved = d->kind;
    if (resolved == CHD_DEC_AUTO) {
        resolved = (system == chd::System::NTSC) ? CHD_DEC_NTSC_2D
                 : (system == chd::System::SECAM) ? CHD_DEC_SECAM : CHD_DEC_PAL_2D;
    }

    int32_t padding = 1;
    if (auto it = d->i32.find("padding_multiple"); it != d->i32.end()) {
        padding = it->second;
        if (padding < 1 || padding > 1024)
            return fail(CHD_E_INVALID_ARG, fn + ": padding_multiple must be between 1 (no padding) and 1024");
    }
    bool y4m = false;
    if (auto it = d->boolean.find("output_y4m_headers"); it != d->boolean.end()) y4m = it->second;

    chd_pixel_format_t fmt = CHD_PIXEL_YUV444P16;
    int fmt_ok = 1;
    if (auto it = d->str.find("output_format"); it != d->str.end()) {
        fmt_ok = parse_format(it->second, fmt);
        if (fmt_ok < 0) return fail(CHD_E_INVALID_ARG, fn + ": unknown output_format \"" + it->second + "\"");
    }
    int clamp = CHD_CLAMP_NONE;
    if (auto it = d->str.find("output_clamp"); it != d->str.end()) {
        clamp = parse_clamp(it->second);
        if (clamp < 0) return fail(CHD_E_INVALID_ARG, fn + ": unknown output_clamp \"" + it->second + "\"");
    }
    std::string cdp, bsp;
    if (auto it = d->str.find("color_difference_precision"); it != d->str.end()) {
        cdp = it->second;
        if (cdp != "classic" && cdp != "modern")
            return fail(CHD_E_INVALID_ARG, fn + ": unknown color_difference_precision \"" + cdp + "\"");
    }
    if (auto it = d->str.find("broadcast_scaling_precision"); it != d->str.end()) {
        bsp = it->second;
        if (bsp != "classic" && bsp != "modern" && bsp != "scientific")
            return fail(CHD_E_INVALID_ARG, fn + ": unknown broadcast_scaling_precision \"" + bsp + "\"");
    }
    {
        const bool is440 = fmt == CHD_PIXEL_YUV440P16 || fmt == CHD_PIXEL_YUV440PS;
        const bool lumaOnly = fmt == CHD_PIXEL_GRAY16 || fmt == CHD_PIXEL_GRAYS;
        if (system == chd::System::SECAM && !is440 && !lumaOnly)
            return fail(CHD_E_INVALID_ARG, fn + ": SECAM sources decode to 4:4:0 chroma or luma-only");
        if (is440 && system != chd::System::SECAM)
            return fail(CHD_E_INVALID_ARG, fn + ": 4:4:0 output is produced only by line-sequential (SECAM) decodes");
        if (is440 && y4m) return fail(CHD_E_INVALID_ARG, fn + ": output_y4m_headers does not support 4:4:0 output");
    }

    // crop overrides and validation (c_api_decoder.cpp:96-143, 829-837)
    chd::VideoParams vp = meta.vp;
    if (auto it = d->i32.find("first_active_frame_line"); it != d->i32.end()) vp.firstActiveFrameLine = it->second;
    if (auto it = d->i32.find("last_active_frame_line"); it != d->i32.end()) vp.lastActiveFrameLine = it->second;
    if (auto it = d->i32.find("first_active_sample"); it != d->i32.end()) vp.activeVideoStart = it->second;
    if (auto it = d->i32.find("last_active_sample"); it != d->i32.end()) vp.activeVideoEnd = it->second + 1;
    if (vp.activeVideoStart < 0 || vp.activeVideoEnd > vp.fieldWidth || vp.activeVideoEnd <= vp.activeVideoStart)
        return fail(CHD_E_INVALID_ARG, fn + ": active sample crop must satisfy 0 <= first_active_sample "
                                            "<= last_active_sample < field_width");
    if (vp.firstActiveFrameLine < 0 || vp.lastActiveFrameLine >= (vp.fieldHeight * 2) - 1 ||
        vp.lastActiveFrameLine < vp.firstActiveFrameLine)
        return fail(CHD_E_INVALID_ARG, fn + ": active frame-line crop must satisfy 0 <= first_active_frame_line "
                                            "<= last_active_frame_line < (2 * field_height) - 1");
    bool reverse = false;
    if (auto it = d->boolean.find("reverse_field_order"); it != d->boolean.end()) reverse = it->second;

    // OutputWriter::updateConfiguration (output_writer.cpp:195-239)
    const int32_t aw = vp.activeVideoEnd - vp.activeVideoStart;
    const int32_t ah = vp.lastActiveFrameLine - vp.firstActiveFrameLine + 1;
    int32_t ow = aw, oh = ah, padL = 0, padR = 0, padT = 0, padB = 0;
    if (padding > 1) {
        while (ow % padding != 0) {
            if (ow % 2 == 0) padR++; else padL++;
            ow = padL + aw + padR;
        }
        while (oh % padding != 0) {
            if (oh % 2 == 0) padB++; else padT++;
            oh = padT + ah + padB;
        }
    }

    d->reverseFieldOrder = reverse;
    d->vp = vp;
    d->dropParams = dropout_params(vp);
    d->format = fmt;
    d->outW = ow;
    d->outH = oh;
    d->padL = padL;
    d->padT = padT;
    d->planes = (fmt == CHD_PIXEL_GRAY16 || fmt == CHD_PIXEL_GRAYS) ? 1 : 3;

    if (d->kind == CHD_DEC_NONE) {
        d->resolved = CHD_DEC_NONE;
        d->lookBehind = d->lookAhead = 0;
        d->threadCount = 0;
        d->committed = true;
        return CHD_OK;
    }

    if (resolved != CHD_DEC_PAL_2D && resolved != CHD_DEC_TRANSFORM_2D && resolved != CHD_DEC_TRANSFORM_3D)
        return fail(CHD_E_DECODER_UNKNOWN, fn + ": unknown or unsupported decoder kind (this build decodes "
                                                "CHD_DEC_PAL_2D and CHD_DEC_TRANSFORM_2D/3D)");
    if (fmt_ok == 0)
        return fail(CHD_E_UNSUPPORTED, fn + ": output_format not produced by this build "
                                            "(yuv444p16, yuv444ps, gray16, grays, rgb48, rgbs are)");
    // PalDecoder::configure (pal_decoder.cpp:38-51)
    if (system != chd::System::PAL && system != chd::System::PAL_M)
        return fail(CHD_E_DECODER_INCOMPATIBLE, fn + ": This decoder is for PAL video sources only");

    int32_t threads = 0;
    if (auto it = d->i32.find("thread_count"); it != d->i32.end()) threads = it->second;
    if (threads <= 0) {
        const unsigned hc = std::thread::hardware_concurrency();
        threads = hc == 0 ? 2 : (int32_t)hc;
    }

    double luma_nr = 0.0;
    if (auto it = d->f64.find("luma_nr_level"); it != d->f64.end()) luma_nr = it->second;
    std::string mask = "threshold";
    if (auto it = d->str.find("transform_mask"); it != d->str.end()) mask = it->second;
    if (mask != "threshold" && mask != "learned")
        return fail(CHD_E_INVALID_ARG, fn + ": unknown transform_mask \"" + mask + "\"");
    d->maskWeights.clear();
    if (mask == "learned") {
        // an explicit weights file must load; without one the packaged
        // weights are used, and if they are absent the threshold stays
        std::string path;
        const auto it = d->str.find("transform_mask_weights");
        const bool explicit_path = it != d->str.end() && !it->second.empty();
        path = explicit_path ? it->second : default_mask_weights();
        if (mask_weights_ok(path)) {
            d->maskWeights = path;
        } else if (explicit_path) {
            return fail(CHD_E_NN_MODEL_LOAD, fn + ": transform_mask_weights \"" + path +
                                                 "\" is not a readable mask weights file");
        } else {
            chd::log_message(CHD_LOG_WARN, fn + ": no learned-mask weights at \"" + path +
                                               "\"; decoding with the threshold mask");
        }
    }

    // PalColour::Configuration + TransformPal::updateConfiguration
    double threshold = 0.4;
    if (auto it = d->f64.find("transform_threshold"); it != d->f64.end()) threshold = it->second;
    std::vector<double> thr;
    if (auto it = d->str.find("transform_thresholds_file"); it != d->str.end() && !it->second.empty()) {
        // registry.cpp:85-92: a file that cannot be opened leaves the
        // thresholds uniform; values are whitespace separated.
        std::ifstream in(it->second);
        if (in.is_open()) {
            double v;
            while (in >> v) thr.push_back(v);
        }
        const size_t want = (resolved == CHD_DEC_TRANSFORM_3D) ? 768 : 80;
        if (resolved != CHD_DEC_PAL_2D && !thr.empty() && thr.size() != want)
            return fail(CHD_E_INVALID_ARG, fn + ": transform_thresholds_file holds " + std::to_string(thr.size()) +
                                               " values, the decoder needs " + std::to_string(want));
    }
    double ecb = 0.0, ecr = 0.0, kr = 0.0, kb = 0.0;
    color_scales(cdp, bsp, ecb, ecr, kr, kb);

    chd_tpal_params_t &p = d->params;
    std::memset(&p, 0, sizeof(p));
    p.kind = resolved;
    p.field_width = vp.fieldWidth;
    p.field_height = vp.fieldHeight;
    p.active_video_start = vp.activeVideoStart;
    p.active_video_end = vp.activeVideoEnd;
    p.first_active_frame_line = vp.firstActiveFrameLine;
    p.last_active_frame_line = vp.lastActiveFrameLine;
    p.colour_burst_start = vp.colourBurstStart;
    p.colour_burst_end = vp.colourBurstEnd;
    p.black_16b_ire = vp.black16bIre;
    p.white_16b_ire = vp.white16bIre;
    p.fsc_hz = vp.fSC;
    p.sample_rate_hz = vp.sampleRate;
    p.threshold = threshold;
    d->thresholds = thr;
    p.thresholds = d->thresholds.empty() ? nullptr : d->thresholds.data();
    p.chroma_gain = 1.0;
    p.chroma_phase_deg = 0.0;
    if (auto it = d->f64.find("chroma_gain"); it != d->f64.end()) p.chroma_gain = it->second;
    if (auto it = d->f64.find("chroma_phase_deg"); it != d->f64.end()) p.chroma_phase_deg = it->second;
    p.ecb_from_u = ecb;
    p.ecr_from_v = ecr;
    p.pixel_format = fmt;
    p.clamp = clamp;
    p.out_width = ow;
    p.out_height = oh;
    p.pad_left = padL;
    p.pad_top = padT;
    p.max_batch_frames = 1;
    p.device = 0;
    p.kr = kr;
    p.kb = kb;
    p.luma_nr_level = luma_nr;

    d->resolved = resolved;
    d->lookBehind = (resolved == CHD_DEC_TRANSFORM_3D) ? 2 : 0;  // (HALFZTILE + 1) / 2
    d->lookAhead = (resolved == CHD_DEC_TRANSFORM_3D) ? 4 : 0;   // (ZTILE - 1 + 1) / 2
    d->threadCount = threads;

    d->framePool = std::make_shared<FramePool>((size_t)d->planes * d->outW * d->outH *
                                                   (d->format == CHD_PIXEL_YUV444PS || d->format == CHD_PIXEL_GRAYS ||
                                                            d->format == CHD_PIXEL_RGBS
                                                        ? sizeof(float)
                                                        : sizeof(uint16_t)),
                                               kBatch);
    // Build worker 0 now so a missing GPU fails here, not at first decode.
    Worker *w0 = get_worker(d, 0);
    chd_status_t st = worker_init(d, *w0, "chd_decoder_commit");
    if (st != CHD_OK) {
        std::lock_guard<std::mutex> lk(d->workersMu);
        d->workers.clear();
        return st;
    }
    d->committed = true;
    return CHD_OK;
}

chd_status_t chd_decoder_get_output_info(const chd_decoder_t *d, chd_output_info_t *out)
{
    if (d == nullptr || out == nullptr)
        return fail(CHD_E_INVALID_ARG, "chromadec: chd_decoder_get_output_info: null argument");
    if (!d->committed) return fail(CHD_E_INVALID_ARG, "chd_decoder_get_output_info: chd_decoder_commit not called");
    out->format = d->format;
    out->width = d->outW;
    out->height = d->outH;
    out->num_planes = d->planes;
    out->num_frames = d->video->meta.numberOfFrames(d->firstFieldFirst());
    return CHD_OK;
}

chd_status_t chd_decode_frame(chd_decoder_t *d, int64_t frame_index, chd_frame_t **out)
{
    if (d == nullptr || out == nullptr) return fail(CHD_E_INVALID_ARG, "chromadec: chd_decode_frame: null argument");
    *out = nullptr;
    if (!d->committed) return fail(CHD_E_INVALID_ARG, "chd_decode_frame: chd_decoder_commit not called");
    if (d->resolved == CHD_DEC_NONE)
        return fail(CHD_E_DECODER_INCOMPATIBLE, "chd_decode_frame: decoder kind is CHD_DEC_NONE (geometry/dropout only)");
    Worker *w = get_worker(d, 0);
    std::lock_guard<std::mutex> lk(w->mu);
    return decode_locked(d, *w, frame_index, out);
}

chd_status_t chd_decode_frames_async(chd_decoder_t *d, const int64_t *indices, size_t n, chd_frame_done_cb cb,
                                      void *user, chd_cancel_t *cancel)
{
    if (d == nullptr || cb == nullptr) return fail(CHD_E_INVALID_ARG, "chromadec: chd_decode_frames_async: null argument");
    if (n > 0 && indices == nullptr)
        return fail(CHD_E_INVALID_ARG, "chromadec: chd_decode_frames_async: null argument (indices)");
    if (!d->committed) return fail(CHD_E_INVALID_ARG, "chd_decode_frames_async: chd_decoder_commit not called");
    if (d->resolved == CHD_DEC_NONE)
        return fail(CHD_E_DECODER_INCOMPATIBLE,
                    "chd_decode_frames_async: decoder kind is CHD_DEC_NONE (geometry/dropout only)");
    if (n == 0) return CHD_OK;
    const size_t W = std::min<size_t>({n, (size_t)std::max(1, d->threadCount), (size_t)kMaxGpuWorkers});
    // Workers take runs of consecutive list positions: neighbouring frames of
    // a run share most of their window fields in that worker's field cache
    // (a 3D window spans 7 frames), where single positions would reload it.
    const size_t run = std::max<size_t>(1, std::min<size_t>(64, (n + W - 1) / W));
    std::atomic<size_t> next{0};
    std::vector<std::thread> threads;
    threads.reserve(W);
    for (size_t wi = 0; wi < W; ++wi) {
        // async workers are 1..W: worker 0 is chd_decode_frame's, so a callback
        // that decodes a frame synchronously does not wait on its own worker
        Worker *w = get_worker(d, wi + 1);
        threads.emplace_back([d, w, indices, n, run, cb, user, cancel, &next]() {
            // two submissions in flight (slots alternate), each of up to
            // kBatch consecutive frame indices; a submission's frames are
            // delivered after the next one is submitted, in list order
            Submission sub[2];
            int k = 0;
            auto deliver = [&](int s) {
                finish_frames(*w, s, sub[s], [&](chd_status_t st, int64_t idx, chd_frame_t *f) { cb(user, st, idx, f); });
            };
            std::lock_guard<std::mutex> lk(w->mu);
            for (;;) {
                const size_t i0 = next.fetch_add(run, std::memory_order_relaxed);
                if (i0 >= n) break;
                const size_t end = std::min(n, i0 + run);
                for (size_t i = i0; i < end;) {
                    int nb = 1;
                    while (nb < kBatch && i + nb < end && indices[i + nb] == indices[i + nb - 1] + 1) ++nb;
                    const int s = k & 1;
                    if (sub[s].busy) deliver(s);
                    if (cancel != nullptr && cancel->requested.load(std::memory_order_acquire)) {
                        if (sub[s ^ 1].busy) deliver(s ^ 1);   // keep list order
                        for (int j = 0; j < nb; ++j) cb(user, CHD_E_CANCELLED, indices[i + j], nullptr);
                    } else {
                        submit_frames(d, *w, indices + i, nb, s, sub[s]);
                        ++k;
                    }
                    i += nb;
                }
            }
            // the last submissions, oldest first
            if (sub[k & 1].busy) deliver(k & 1);
            if (sub[(k + 1) & 1].busy) deliver((k + 1) & 1);
        });
    }
    for (auto &t : threads) t.join();
    return CHD_OK;
}

// ---- frames (c_api_frame.cpp) ----------------------------------------------

chd_status_t chd_frame_get_info(const chd_frame_t *f, chd_frame_info_t *out)
{
    if (f == nullptr || out == nullptr) return fail(CHD_E_INVALID_ARG, "chromadec: chd_frame_get_info: null argument");
    *out = f->info;
    return CHD_OK;
}

chd_status_t chd_frame_get_plane_info(const chd_frame_t *f, chd_plane_t p, chd_plane_info_t *out)
{
    if (f == nullptr || out == nullptr) return fail(CHD_E_INVALID_ARG, "chromadec: chd_frame_get_plane_info: null argument");
    bool valid = false;
    switch (f->format) {
    case CHD_PIXEL_YUV444P16:
    case CHD_PIXEL_YUV444PS:
        valid = p == CHD_PLANE_Y || p == CHD_PLANE_CB || p == CHD_PLANE_CR;
        break;
    case CHD_PIXEL_GRAY16:
    case CHD_PIXEL_GRAYS: valid = p == CHD_PLANE_Y; break;
    case CHD_PIXEL_RGB48:
    case CHD_PIXEL_RGBS: valid = p == CHD_PLANE_R || p == CHD_PLANE_G || p == CHD_PLANE_B; break;
    default: break;
    }
    if (!valid) return fail(CHD_E_INVALID_ARG, "chd_frame_get_plane_info: plane not valid for the frame's pixel format");
    out->width = f->width;
    out->height = f->height;
    out->first_frame_row = 0;
    return CHD_OK;
}

chd_status_t chd_frame_get_plane(const chd_frame_t *f, chd_plane_t p, const void **out_data, ptrdiff_t *out_stride)
{
    if (f == nullptr || out_data == nullptr || out_stride == nullptr)
        return fail(CHD_E_INVALID_ARG, "chromadec: chd_frame_get_plane: null argument");
    *out_data = nullptr;
    *out_stride = 0;
    const size_t ps = (size_t)f->width * f->height;
    switch (f->format) {
    case CHD_PIXEL_YUV444P16:
        if (p != CHD_PLANE_Y && p != CHD_PLANE_CB && p != CHD_PLANE_CR)
            return fail(CHD_E_INVALID_ARG, "chd_frame_get_plane: plane not valid for YUV444P16");
        *out_data = f->u16() + ps * (size_t)p;
        *out_stride = (ptrdiff_t)f->width * (ptrdiff_t)sizeof(uint16_t);
        return CHD_OK;
    case CHD_PIXEL_GRAY16:
        if (p != CHD_PLANE_Y) return fail(CHD_E_INVALID_ARG, "chd_frame_get_plane: GRAY16 only has CHD_PLANE_Y");
        *out_data = f->u16();
        *out_stride = (ptrdiff_t)f->width * (ptrdiff_t)sizeof(uint16_t);
        return CHD_OK;
    case CHD_PIXEL_RGB48:
        // interleaved R'G'B': the three planes share the buffer at sample
        // offsets 0/1/2 with the row pitch as stride (c_api_frame.cpp:161-179)
        if (p != CHD_PLANE_R && p != CHD_PLANE_G && p != CHD_PLANE_B)
            return fail(CHD_E_INVALID_ARG, "chd_frame_get_plane: plane not valid for RGB48");
        *out_data = f->u16() + (size_t)(p - CHD_PLANE_R);
        *out_stride = (ptrdiff_t)f->width * 3 * (ptrdiff_t)sizeof(uint16_t);
        return CHD_OK;
    case CHD_PIXEL_YUV444PS:
    case CHD_PIXEL_GRAYS:
    case CHD_PIXEL_RGBS:
        return fail(CHD_E_INVALID_ARG, "chd_frame_get_plane: use chd_frame_get_plane_float for float pixel formats");
    default: break;
    }
    return fail(CHD_E_INTERNAL, "chd_frame_get_plane: unknown pixel format");
}

chd_status_t chd_frame_get_plane_float(const chd_frame_t *f, chd_plane_t p, const float **out_data,
                                        ptrdiff_t *out_stride)
{
    if (f == nullptr || out_data == nullptr || out_stride == nullptr)
        return fail(CHD_E_INVALID_ARG, "chromadec: chd_frame_get_plane_float: null argument");
    if (f->format != CHD_PIXEL_YUV444PS && f->format != CHD_PIXEL_GRAYS && f->format != CHD_PIXEL_RGBS)
        return fail(CHD_E_INVALID_ARG, "chd_frame_get_plane_float: frame is not a float pixel format");
    if (f->format == CHD_PIXEL_GRAYS && p != CHD_PLANE_Y)
        return fail(CHD_E_INVALID_ARG, "chd_frame_get_plane_float: CHD_PIXEL_GRAYS only has CHD_PLANE_Y");
    size_t idx;
    if (f->format == CHD_PIXEL_RGBS) {
        if (p != CHD_PLANE_R && p != CHD_PLANE_G && p != CHD_PLANE_B)
            return fail(CHD_E_INVALID_ARG, "chd_frame_get_plane_float: CHD_PIXEL_RGBS plane must be R, G, or B");
        idx = (size_t)(p - CHD_PLANE_R);
    } else {
        if (p != CHD_PLANE_Y && p != CHD_PLANE_CB && p != CHD_PLANE_CR)
            return fail(CHD_E_INVALID_ARG, "chd_frame_get_plane_float: Y'CbCr float frame plane must be Y, Cb, or Cr");
        idx = (size_t)p;
    }
    *out_data = f->f32() + (size_t)f->width * f->height * idx;
    *out_stride = (ptrdiff_t)f->width * (ptrdiff_t)sizeof(float);
    return CHD_OK;
}

void chd_frame_free(chd_frame_t *f) { delete f; }

chd_status_t chd_decoder_set_dropout(chd_decoder_t *d, const chd_dropout_opts_t *opts)
{
    if (d == nullptr || opts == nullptr) return fail(CHD_E_INVALID_ARG, "chd_decoder_set_dropout: null argument");
    std::lock_guard<std::mutex> lk(d->dropMu);
    d->dropoutOpts = *opts;
    d->dropoutSet = true;
    d->dropoutRevision += 1;
    d->corrected.clear();
    d->correctedOrder.clear();
    return CHD_OK;
}

chd_status_t chd_decoder_get_last_dropout_stats(const chd_decoder_t *d, chd_dropout_stats_t *out)
{
    if (d == nullptr || out == nullptr)
        return fail(CHD_E_INVALID_ARG, "chd_decoder_get_last_dropout_stats: null argument");
    std::lock_guard<std::mutex> lk(const_cast<chd_decoder_t *>(d)->dropMu);
    *out = d->lastStats;
    return CHD_OK;
}

chd_status_t chd_video_correct_dropouts(const chd_video_t *video, int64_t frame_index, const chd_dropout_opts_t *opts,
                                        uint16_t *first_field, uint16_t *second_field, chd_dropout_stats_t *stats)
{
    if (video == nullptr || opts == nullptr || first_field == nullptr || second_field == nullptr)
        return fail(CHD_E_INVALID_ARG, "chd_video_correct_dropouts: null argument");
    const chd::Metadata &meta = video->meta;
    const bool fff = meta.isFirstFieldFirst;
    if (frame_index < 0 || frame_index >= meta.numberOfFrames(fff))
        return fail(CHD_E_OUT_OF_RANGE, "chd_video_correct_dropouts: frame_index out of range");
    const int32_t f1 = meta.fieldNumber((int32_t)frame_index + 1, false, fff);
    const int32_t f2 = meta.fieldNumber((int32_t)frame_index + 1, true, fff);
    chd::dropout::Stats st;
    if (opts->enabled)
        chd::dropout::correct_frame(dropout_params(meta.vp), first_field, second_field, runs_of(meta.field(f1).dropOuts),
                                    runs_of(meta.field(f2).dropOuts), opts->overcorrect != 0,
                                    opts->intra_field_only != 0, &st);
    if (stats) {
        stats->corrected = st.corrected;
        stats->failed = st.failed;
        stats->total_distance = st.total_distance;
    }
    return CHD_OK;
}

}  // extern "C"
