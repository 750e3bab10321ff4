// Composite sources: ld-decode `.tbc` + `.tbc.json`, metadata, chd_video_*.
//
// Reference behaviour followed:
//   src/metadata/ld_metadata_json.cpp:65-116,233-318,324-360  sidecar keys
//   src/metadata/core.cpp:40-110,742-811                       system defaults
//   src/metadata/core.cpp:1092-1190                            frame pairing
//   src/abi/c_api_video.cpp:511-584,650-790                    open / info / lines
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "common.h"
#include "handles.h"
#include "json.h"

namespace chd {

// ---------------------------------------------------------------------------
// Metadata
// ---------------------------------------------------------------------------

const FieldMeta &Metadata::field(int32_t number) const
{
    static const FieldMeta empty{};
    if (number < 1 || number > numberOfFields()) return empty;
    return fields[number - 1];
}

int32_t Metadata::numberOfFrames(bool firstFieldFirst) const
{
    int32_t frameOffset = 0;
    if (firstFieldFirst) {
        if (!field(1).isFirstField) frameOffset = 1;
    } else {
        if (field(1).isFirstField) frameOffset = 1;
    }
    return (numberOfFields() / 2) - frameOffset;
}

int32_t Metadata::fieldNumber(int32_t frame, bool second, bool firstFieldFirst) const
{
    if (frame < 1) return -1;
    int32_t firstNo, secondNo;
    if (firstFieldFirst) {
        firstNo = frame * 2 - 1;
        secondNo = firstNo + 1;
    } else {
        secondNo = frame * 2 - 1;
        firstNo = secondNo + 1;
    }
    // advance until the first field really is a first field
    while (!field(firstNo).isFirstField) {
        firstNo++;
        secondNo++;
        if (firstNo > numberOfFields() || secondNo > numberOfFields()) {
            log_message(CHD_LOG_ERROR, "frame " + std::to_string(frame) + ": no field marked isFirstField before "
                                       "the end of the metadata");
            firstNo = secondNo = -1;
            break;
        }
    }
    if (firstNo > numberOfFields() || secondNo > numberOfFields()) {
        log_message(CHD_LOG_ERROR, "frame " + std::to_string(frame) + ": field number beyond the last field");
        firstNo = secondNo = -1;
    }
    if (firstNo == -1) return -1;
    // both fields of the pair marked first: a broken capture, decoded as is
    if (field(secondNo).isFirstField)
        log_message(CHD_LOG_WARN, "frame " + std::to_string(frame) + ": fields " + std::to_string(firstNo) + " and " +
                                      std::to_string(secondNo) + " are both marked isFirstField");
    return second ? secondNo : firstNo;
}

namespace {

std::string normalise_system(std::string n)
{
    while (!n.empty() && std::isspace((unsigned char)n.front())) n.erase(n.begin());
    while (!n.empty() && std::isspace((unsigned char)n.back())) n.pop_back();
    for (auto &c : n) c = (c == '-' || c == ' ') ? '_' : (char)std::toupper((unsigned char)c);
    return n;
}

}  // namespace

bool parse_system(const std::string &name, System &out)
{
    const std::string n = normalise_system(name);
    if (n == "PAL") { out = System::PAL; return true; }
    if (n == "NTSC") { out = System::NTSC; return true; }
    if (n == "SECAM") { out = System::SECAM; return true; }
    if (n == "PAL_M" || n == "MPAL" || n == "PALM" || n == "N_LINHA" || n == "NLINHA" || n == "PAL_N_LINHA" ||
        n == "PAL_NLINHA") {
        out = System::PAL_M;
        return true;
    }
    return false;
}

void apply_system_defaults(VideoParams &vp)
{
    switch (vp.system) {
    case System::PAL:
        vp.fSC = (283.75 * 15625) + 25;
        vp.firstActiveFrameLine = 44;
        vp.lastActiveFrameLine = 619;
        break;
    case System::SECAM:
        vp.fSC = 4286000.0;
        vp.firstActiveFrameLine = 44;
        vp.lastActiveFrameLine = 619;
        break;
    case System::NTSC:
        vp.fSC = 315.0e6 / 88.0;
        vp.firstActiveFrameLine = 39;
        vp.lastActiveFrameLine = 524;
        break;
    case System::PAL_M:
        vp.fSC = 5.0e6 * (63.0 / 88.0) * (909.0 / 910.0);
        vp.firstActiveFrameLine = 39;
        vp.lastActiveFrameLine = 524;
        break;
    }
}

namespace {

bool read_int_array(const json::Value *v, std::vector<int32_t> &out)
{
    out.clear();
    if (v == nullptr) return true;
    if (v->type != json::Value::Array) return false;
    for (const auto &e : v->a) {
        if (!e.is_number()) return false;
        out.push_back((int32_t)e.n);
    }
    return true;
}

}  // namespace

chd_status_t read_ld_json(const std::string &path, Metadata &meta, std::string &err)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        err = "cannot open sidecar " + path;
        return CHD_E_FILE_NOT_FOUND;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    json::Value doc;
    if (!json::parse(ss.str(), doc, err)) {
        err = "sidecar " + path + " does not parse: " + err;
        return CHD_E_METADATA_CORRUPT;
    }
    if (doc.type != json::Value::Object) {
        err = "sidecar root is not an object";
        return CHD_E_METADATA_CORRUPT;
    }
    const json::Value *vpj = doc.get("videoParameters");
    if (vpj == nullptr || vpj->type != json::Value::Object) {
        err = "JSON file invalid: videoParameters object is not defined";
        return CHD_E_METADATA_CORRUPT;
    }
    VideoParams vp;
    auto gi = [&](const char *k, int32_t dflt) {
        const json::Value *x = vpj->get(k);
        return x ? x->as_int(dflt) : dflt;
    };
    vp.numberOfSequentialFields = gi("numberOfSequentialFields", -1);
    vp.colourBurstStart = gi("colourBurstStart", -1);
    vp.colourBurstEnd = gi("colourBurstEnd", -1);
    vp.activeVideoStart = gi("activeVideoStart", -1);
    vp.activeVideoEnd = gi("activeVideoEnd", -1);
    vp.white16bIre = gi("white16bIre", -1);
    vp.black16bIre = gi("black16bIre", -1);
    vp.blanking16bIre = gi("blanking16bIre", -1);
    vp.fieldWidth = gi("fieldWidth", -1);
    vp.fieldHeight = gi("fieldHeight", -1);
    if (const json::Value *x = vpj->get("sampleRate")) vp.sampleRate = x->as_double(-1.0);
    if (const json::Value *x = vpj->get("isSubcarrierLocked")) vp.isSubcarrierLocked = x->as_bool(false);
    if (const json::Value *x = vpj->get("isWidescreen")) vp.isWidescreen = x->as_bool(false);
    bool isSourcePal = false;
    if (const json::Value *x = vpj->get("isSourcePal")) isSourcePal = x->as_bool(false);
    const json::Value *sys = vpj->get("system");
    if (sys == nullptr || sys->type != json::Value::String || sys->s.empty()) {
        vp.system = isSourcePal ? (vp.fieldHeight < 300 ? System::PAL_M : System::PAL) : System::NTSC;
    } else if (!parse_system(sys->s, vp.system)) {
        err = "unknown value for videoParameters.system";
        return CHD_E_METADATA_CORRUPT;
    }
    if (vp.blanking16bIre == -1) vp.blanking16bIre = vp.black16bIre;

    std::vector<FieldMeta> fields;
    if (const json::Value *fl = doc.get("fields")) {
        if (fl->type != json::Value::Array) {
            err = "fields is not an array";
            return CHD_E_METADATA_CORRUPT;
        }
        fields.reserve(fl->a.size());
        for (const auto &fj : fl->a) {
            FieldMeta fm;
            if (const json::Value *x = fj.get("isFirstField")) fm.isFirstField = x->as_bool(false);
            if (const json::Value *d = fj.get("dropOuts")) {
                if (!read_int_array(d->get("startx"), fm.dropOuts.startx) ||
                    !read_int_array(d->get("endx"), fm.dropOuts.endx) ||
                    !read_int_array(d->get("fieldLine"), fm.dropOuts.fieldLine) ||
                    fm.dropOuts.startx.size() != fm.dropOuts.endx.size() ||
                    fm.dropOuts.startx.size() != fm.dropOuts.fieldLine.size()) {
                    err = "dropout array sizes do not match";
                    return CHD_E_METADATA_CORRUPT;
                }
            }
            fields.push_back(std::move(fm));
        }
    }
    if ((size_t)vp.numberOfSequentialFields != fields.size()) {
        err = "JSON file invalid: numberOfSequentialFields does not match fields array";
        return CHD_E_METADATA_CORRUPT;
    }
    apply_system_defaults(vp);
    meta.vp = vp;
    meta.fields = std::move(fields);
    meta.isFirstFieldFirst = true;
    return CHD_OK;
}

// ---------------------------------------------------------------------------
// TBC reader
// ---------------------------------------------------------------------------

TbcFile::~TbcFile()
{
    if (fd >= 0) ::close(fd);
}

bool TbcFile::open(const std::string &path, int64_t samples_per_field, std::string &err)
{
    fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd < 0) {
        err = "cannot open " + path + ": " + std::strerror(errno);
        return false;
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
        err = "cannot stat " + path;
        return false;
    }
    field_samples = samples_per_field;
    available_fields = samples_per_field > 0 ? (int32_t)(st.st_size / (samples_per_field * 2)) : 0;
    return true;
}

bool TbcFile::read_field(int32_t number, uint16_t *dst) const
{
    if (number < 1 || number > available_fields) return false;
    const int64_t bytes = field_samples * 2;
    int64_t done = 0;
    const off_t base = (off_t)(number - 1) * bytes;
    while (done < bytes) {
        const ssize_t r = ::pread(fd, (char *)dst + done, bytes - done, base + done);
        if (r <= 0) return false;
        done += r;
    }
    return true;
}

}  // namespace chd

// ---------------------------------------------------------------------------
// ABI
// ---------------------------------------------------------------------------

namespace {

bool file_exists(const std::string &p)
{
    struct stat st;
    return ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

bool ends_with(const std::string &s, const char *suf)
{
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

chd_video_standard_t to_abi(chd::System s)
{
    switch (s) {
    case chd::System::PAL: return CHD_STD_PAL;
    case chd::System::NTSC: return CHD_STD_NTSC;
    case chd::System::PAL_M: return CHD_STD_PAL_M;
    case chd::System::SECAM: return CHD_STD_SECAM;
    }
    return CHD_STD_UNKNOWN;
}

}  // namespace

extern "C" {

chd_status_t chd_init(void) { return CHD_OK; }
void chd_shutdown(void) {}

chd_status_t chd_video_open_composite(const char *path, const char *metadata_path_or_null,
                                      const chd_video_params_t *override_or_null, chd_video_t **out)
{
    const std::string fn = "chd_video_open_composite";
    if (path == nullptr || out == nullptr) return chd::fail(CHD_E_INVALID_ARG, fn + ": null argument");
    const std::string p = path;
    if (!file_exists(p)) return chd::fail(CHD_E_FILE_NOT_FOUND, fn + ": file does not exist: " + p);
    std::string side;
    if (metadata_path_or_null != nullptr) {
        side = metadata_path_or_null;
        if (!file_exists(side)) return chd::fail(CHD_E_FILE_NOT_FOUND, fn + ": sidecar file does not exist: " + side);
    } else if (file_exists(p + ".db")) {
        // resolveSidecarFlavour (c_api_video.cpp:304-321): `.db` before `.json`
        side = p + ".db";
    } else if (file_exists(p + ".json")) {
        side = p + ".json";
    } else {
        return chd::fail(CHD_E_METADATA_MISSING, fn + ": no `.db` or `.json` sidecar next to " + p);
    }
    if (ends_with(side, ".meta"))
        return chd::fail(CHD_E_FORMAT_UNSUPPORTED, fn + ": CVBS `.meta` sidecars are not supported by this build");
    auto v = std::make_unique<chd_video>();
    v->path = p;
    std::string err;
    // LdDecodeMetaData::read (core.cpp:335-350): `.json` is JSON, anything
    // else SQLite
    std::string lower = side;
    for (char &c : lower) c = (char)std::tolower((unsigned char)c);
    const chd_status_t st = ends_with(lower, ".json") ? chd::read_ld_json(side, v->meta, err)
                                                      : chd::read_ld_sqlite(side, v->meta, err);
    if (st != CHD_OK) return chd::fail(st, fn + ": " + err);
    chd::VideoParams &vp = v->meta.vp;
    if (override_or_null != nullptr) {
        const chd_video_params_t &o = *override_or_null;
        if (o.standard != CHD_STD_UNKNOWN && o.standard != to_abi(vp.system)) {
            const bool declared625 = (o.standard == CHD_STD_PAL || o.standard == CHD_STD_SECAM);
            const bool current625 = (vp.system == chd::System::PAL || vp.system == chd::System::SECAM);
            if (declared625 != current625)
                return chd::fail(CHD_E_INVALID_ARG, fn + ": declared standard's line standard does not match the capture's");
            // LdDecodeMetaData::overrideVideoSystem (core.cpp:760-769): new
            // system, its subcarrier, the same (re-validated) active lines.
            vp.system = o.standard == CHD_STD_PAL ? chd::System::PAL
                      : o.standard == CHD_STD_SECAM ? chd::System::SECAM
                      : o.standard == CHD_STD_NTSC ? chd::System::NTSC : chd::System::PAL_M;
            switch (vp.system) {
            case chd::System::PAL: vp.fSC = (283.75 * 15625) + 25; break;
            case chd::System::SECAM: vp.fSC = 4286000.0; break;
            case chd::System::NTSC: vp.fSC = 315.0e6 / 88.0; break;
            case chd::System::PAL_M: vp.fSC = 5.0e6 * (63.0 / 88.0) * (909.0 / 910.0); break;
            }
        }
        if (o.is_subcarrier_locked) vp.isSubcarrierLocked = true;
        v->meta.isFirstFieldFirst = (o.is_second_field_first == 0);
    }
    if (vp.fieldWidth <= 0 || vp.fieldHeight <= 0)
        return chd::fail(CHD_E_METADATA_CORRUPT, fn + ": sidecar lacks the field geometry");
    if (!v->tbc.open(p, (int64_t)vp.fieldWidth * vp.fieldHeight, err)) return chd::fail(CHD_E_IO, fn + ": " + err);
    *out = v.release();
    return CHD_OK;
}

void chd_video_free(chd_video_t *v) { delete v; }

chd_status_t chd_video_get_info(const chd_video_t *v, chd_video_info_t *out)
{
    if (v == nullptr || out == nullptr) return chd::fail(CHD_E_INVALID_ARG, "chd_video_get_info: null argument");
    const chd::VideoParams &vp = v->meta.vp;
    std::memset(out, 0, sizeof(*out));
    out->standard = to_abi(vp.system);
    out->encoding = CHD_ENC_CVBS_U16_4FSC;
    out->signal_state = CHD_SIG_STANDARD_TBC_LOCKED;
    out->layout = CHD_FRAME_LAYOUT_FIELD_RASTER;
    out->field_width = vp.fieldWidth;
    out->field_height = vp.fieldHeight;
    switch (vp.system) {
    case chd::System::PAL:
    case chd::System::SECAM: out->samples_per_frame = 1135 * 625 + 4; break;
    case chd::System::NTSC: out->samples_per_frame = 910 * 525; break;
    case chd::System::PAL_M: out->samples_per_frame = 909 * 525; break;
    }
    out->sample_rate_hz = vp.sampleRate;
    out->fsc_hz = vp.fSC;
    out->first_active_sample = vp.activeVideoStart;
    out->last_active_sample = vp.activeVideoEnd - 1;
    out->first_active_frame_line = vp.firstActiveFrameLine;
    out->last_active_frame_line = vp.lastActiveFrameLine;
    out->black_16b_ire = vp.black16bIre;
    out->white_16b_ire = vp.white16bIre;
    out->blanking_16b_ire = vp.blanking16bIre;
    out->num_frames = v->meta.numberOfFrames(v->meta.isFirstFieldFirst);
    out->is_first_field_first = v->meta.isFirstFieldFirst ? 1 : 0;
    out->is_widescreen = vp.isWidescreen ? 1 : 0;
    out->is_subcarrier_locked = vp.isSubcarrierLocked ? 1 : 0;
    return CHD_OK;
}

chd_status_t chd_video_frame_line_to_signal_line(const chd_video_t *v, int32_t frame_line, int32_t *out)
{
    if (v == nullptr || out == nullptr)
        return chd::fail(CHD_E_INVALID_ARG, "chd_video_frame_line_to_signal_line: null argument");
    const int32_t H = v->meta.vp.fieldHeight;
    if (frame_line < 0 || frame_line > 2 * H - 2)
        return chd::fail(CHD_E_OUT_OF_RANGE,
                         "chd_video_frame_line_to_signal_line: frame_line out of range [0, (2 * field_height) - 2]");
    *out = (frame_line % 2 == 0) ? (frame_line / 2 + 1) : (H + (frame_line + 1) / 2);
    return CHD_OK;
}

chd_status_t chd_video_signal_line_to_frame_line(const chd_video_t *v, int32_t signal_line, int32_t *out)
{
    if (v == nullptr || out == nullptr)
        return chd::fail(CHD_E_INVALID_ARG, "chd_video_signal_line_to_frame_line: null argument");
    const int32_t H = v->meta.vp.fieldHeight;
    if (signal_line < 1 || signal_line > 2 * H - 1)
        return chd::fail(CHD_E_OUT_OF_RANGE,
                         "chd_video_signal_line_to_frame_line: signal_line out of range [1, (2 * field_height) - 1]");
    *out = (signal_line <= H) ? 2 * (signal_line - 1) : 2 * (signal_line - H) - 1;
    return CHD_OK;
}

chd_status_t chd_cancel_create(chd_cancel_t **out)
{
    if (out == nullptr) return chd::fail(CHD_E_INVALID_ARG, "chd_cancel_create: null argument");
    *out = new chd_cancel();
    return CHD_OK;
}

void chd_cancel_request(chd_cancel_t *c)
{
    if (c) c->requested.store(true, std::memory_order_release);
}

int chd_cancel_is_requested(const chd_cancel_t *c)
{
    return (c != nullptr && c->requested.load(std::memory_order_acquire)) ? 1 : 0;
}

void chd_cancel_free(chd_cancel_t *c) { delete c; }

}  // extern "C"
