// ABI handle definitions and the host-side source model.
//
// Reference counterparts: src/abi/handles.h:49-250 (handles),
// src/metadata/core.h:56-118 (VideoParameters), src/metadata/core.cpp:
// 1092-1160 (frame pairing), src/reader/tbc_source.cpp (field reads).
#pragma once

#include <chromadec/chromadec.h>
#include <chromadec/tpal_device.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace chd {

enum class System { PAL = 0, NTSC = 1, PAL_M = 2, SECAM = 3 };

struct VideoParams {
    int32_t numberOfSequentialFields = -1;
    System system = System::NTSC;
    bool isSubcarrierLocked = false;
    bool isWidescreen = false;
    int32_t colourBurstStart = -1, colourBurstEnd = -1;
    int32_t activeVideoStart = -1, activeVideoEnd = -1;  // half-open
    int32_t white16bIre = -1, black16bIre = -1, blanking16bIre = -1;
    int32_t fieldWidth = -1, fieldHeight = -1;
    double sampleRate = -1.0;
    double fSC = -1.0;
    int32_t firstActiveFrameLine = -1, lastActiveFrameLine = -1;  // inclusive
};

struct DropOuts {
    std::vector<int32_t> startx, endx, fieldLine;  // fieldLine is 1-based
    bool empty() const { return startx.empty(); }
};

struct FieldMeta {
    bool isFirstField = false;
    DropOuts dropOuts;
};

struct Metadata {
    VideoParams vp;
    std::vector<FieldMeta> fields;
    bool isFirstFieldFirst = true;

    int32_t numberOfFields() const { return (int32_t)fields.size(); }
    const FieldMeta &field(int32_t number) const;          // 1-based; blank record if out of range
    int32_t numberOfFrames(bool firstFieldFirst) const;     // core.cpp:1092-1107
    int32_t fieldNumber(int32_t frame, bool second, bool firstFieldFirst) const;  // core.cpp:1117-1179
};

// Reads fields of a headerless little-endian uint16 .tbc file.
struct TbcFile {
    int fd = -1;
    int64_t field_samples = 0;
    int32_t available_fields = 0;
    ~TbcFile();
    bool open(const std::string &path, int64_t samples_per_field, std::string &err);
    bool read_field(int32_t number /*1-based*/, uint16_t *dst) const;
};

chd_status_t read_ld_json(const std::string &path, Metadata &meta, std::string &err);
// `.tbc.db` (ld-decode SQLite) sidecar, through the system libsqlite3 loaded
// at run time (sqlite_sidecar.cpp)
chd_status_t read_ld_sqlite(const std::string &path, Metadata &meta, std::string &err);
// videoParameters.system names (parseVideoSystemName, core.cpp:116) and the per-system
// subcarrier and active-line defaults (core.cpp:40-84)
bool parse_system(const std::string &name, System &out);
void apply_system_defaults(VideoParams &vp);

}  // namespace chd

struct chd_video {
    std::string path;
    chd::Metadata meta;
    chd::TbcFile tbc;
};

struct chd_cancel {
    std::atomic<bool> requested{false};
};

// Pinned host blocks of one frame size, shared by a decoder and the frames
// it returns (which may outlive it).  Decoded frames are downloaded straight
// into a block: one frame, or a batch of consecutive frames in one copy.  A
// block returns to the pool when the last frame cut from it is freed.  Reuse
// avoids a fresh 6 MB mapping, its page faults and a bounce copy per frame.
struct FramePool : std::enable_shared_from_this<FramePool> {
    FramePool(size_t frame_bytes, int batch) : bytes(frame_bytes), batch(batch) {}
    ~FramePool();
    // `n` frames (1 or batch) in one block; null when pinned memory cannot be had
    std::shared_ptr<uint8_t> get(int n);
    const size_t bytes;
    const int batch;
    std::mutex mu;
    std::vector<void *> free_one, free_batch;
};

struct chd_frame {
    chd_frame_info_t info{};
    chd_pixel_format_t format = CHD_PIXEL_YUV444P16;
    int32_t width = 0, height = 0, planes = 0;
    std::shared_ptr<uint8_t> block;   // the pinned block `data` lies in
    void *data = nullptr;     // planes x height x width, float or uint16 by format
    chd_frame() = default;
    chd_frame(const chd_frame &) = delete;
    chd_frame &operator=(const chd_frame &) = delete;
    const float *f32() const { return static_cast<const float *>(data); }
    const uint16_t *u16() const { return static_cast<const uint16_t *>(data); }
};
