// Error detail, status names, logging, version and feature queries.
#include "common.h"

#include <chromadec/version.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <shared_mutex>

namespace chd {
namespace {

thread_local std::string t_last_error;

struct Sink {
    std::shared_mutex mu;           // held shared while a message is delivered
    chd_log_fn fn = nullptr;
    void *user = nullptr;
};
Sink &sink()
{
    static Sink *s = new Sink();    // never destroyed: usable from atexit
    return *s;
}
std::atomic<int> g_level{CHD_LOG_INFO};

void deliver(chd_log_level_t level, chd_log_flags_t flags, const std::string &msg)
{
    if (static_cast<int>(level) < g_level.load(std::memory_order_relaxed)) return;
    Sink &s = sink();
    std::shared_lock<std::shared_mutex> lk(s.mu);
    if (s.fn != nullptr) s.fn(level, flags, msg.c_str(), s.user);
}

void stderr_sink(chd_log_level_t level, chd_log_flags_t, const char *message, void *)
{
    std::fprintf(stderr, "chromadec %s: %s\n", chd_log_level_str(level), message);
}

}  // namespace

void set_last_error(const std::string &msg)
{
    t_last_error = msg;
    deliver(CHD_LOG_ERROR, CHD_LOG_F_RETURNED, msg);
}

void clear_last_error() { t_last_error.clear(); }

void log_message(chd_log_level_t level, const std::string &msg) { deliver(level, 0u, msg); }

}  // namespace chd

extern "C" {

const char *chd_status_str(chd_status_t s)
{
    switch (s) {
    case CHD_OK: return "CHD_OK";
    case CHD_E_INVALID_ARG: return "CHD_E_INVALID_ARG";
    case CHD_E_FILE_NOT_FOUND: return "CHD_E_FILE_NOT_FOUND";
    case CHD_E_IO: return "CHD_E_IO";
    case CHD_E_FORMAT_UNSUPPORTED: return "CHD_E_FORMAT_UNSUPPORTED";
    case CHD_E_METADATA_MISSING: return "CHD_E_METADATA_MISSING";
    case CHD_E_METADATA_CORRUPT: return "CHD_E_METADATA_CORRUPT";
    case CHD_E_PRESET_UNKNOWN: return "CHD_E_PRESET_UNKNOWN";
    case CHD_E_DECODER_UNKNOWN: return "CHD_E_DECODER_UNKNOWN";
    case CHD_E_DECODER_INCOMPATIBLE: return "CHD_E_DECODER_INCOMPATIBLE";
    case CHD_E_NN_MODEL_LOAD: return "CHD_E_NN_MODEL_LOAD";
    case CHD_E_NN_BACKEND_UNAVAILABLE: return "CHD_E_NN_BACKEND_UNAVAILABLE";
    case CHD_E_NN_INFERENCE: return "CHD_E_NN_INFERENCE";
    case CHD_E_OUT_OF_RANGE: return "CHD_E_OUT_OF_RANGE";
    case CHD_E_CANCELLED: return "CHD_E_CANCELLED";
    case CHD_E_INTERNAL: return "CHD_E_INTERNAL";
    case CHD_E_OOM: return "CHD_E_OOM";
    case CHD_E_UNSUPPORTED: return "CHD_E_UNSUPPORTED";
    }
    return "CHD_E_UNKNOWN_STATUS";
}

const char *chd_last_error(void) { return chd::t_last_error.c_str(); }
void chd_clear_last_error(void) { chd::t_last_error.clear(); }

void chd_set_log_callback(chd_log_fn fn_or_null, void *user_data)
{
    chd::Sink &s = chd::sink();
    std::unique_lock<std::shared_mutex> lk(s.mu);  // waits for running deliveries
    s.fn = fn_or_null;
    s.user = user_data;
}

void chd_log_to_stderr(void) { chd_set_log_callback(chd::stderr_sink, nullptr); }

void chd_set_log_level(chd_log_level_t min_level)
{
    int v = static_cast<int>(min_level);
    if (v < CHD_LOG_DEBUG) v = CHD_LOG_DEBUG;
    if (v > CHD_LOG_OFF) v = CHD_LOG_OFF;
    chd::g_level.store(v, std::memory_order_relaxed);
}

chd_log_level_t chd_get_log_level(void) { return static_cast<chd_log_level_t>(chd::g_level.load()); }

int chd_log_is_enabled(chd_log_level_t level)
{
    if (level >= CHD_LOG_OFF || level < CHD_LOG_DEBUG) return 0;
    return static_cast<int>(level) >= chd::g_level.load() ? 1 : 0;
}

const char *chd_log_level_str(chd_log_level_t level)
{
    switch (level) {
    case CHD_LOG_DEBUG: return "DEBUG";
    case CHD_LOG_INFO: return "INFO";
    case CHD_LOG_WARN: return "WARN";
    case CHD_LOG_ERROR: return "ERROR";
    case CHD_LOG_OFF: return "OFF";
    }
    return "UNKNOWN";
}

const char *chd_version_string(void) { return CHROMADEC_VERSION_STRING; }

void chd_version(int *major, int *minor, int *patch)
{
    if (major) *major = CHROMADEC_VERSION_MAJOR;
    if (minor) *minor = CHROMADEC_VERSION_MINOR;
    if (patch) *patch = CHROMADEC_VERSION_PATCH;
}

int chd_has_feature(const char *feature)
{
    if (feature == nullptr) return 0;
    if (std::strcmp(feature, "rocm") == 0) return 1;
    return 0;
}

}  // extern "C"
