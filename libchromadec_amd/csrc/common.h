// Host-side shared state: thread-local last-error detail and the diagnostic
// sink (reference: src/common/error_state.cpp:22-44, src/common/log.h:15-93,
// include/chromadec/log.h).
#pragma once

#include <chromadec/errors.h>
#include <chromadec/log.h>

#include <string>

namespace chd {

// Record `msg` as this thread's chd_last_error() detail and, when a sink is
// installed and ERROR is enabled, log it with CHD_LOG_F_RETURNED.
void set_last_error(const std::string &msg);
void clear_last_error();

// Convenience: set the detail and return `status`.
inline chd_status_t fail(chd_status_t status, const std::string &msg)
{
    set_last_error(msg);
    return status;
}

// Emit a diagnostic that is not tied to a returned status.
void log_message(chd_log_level_t level, const std::string &msg);

}  // namespace chd
