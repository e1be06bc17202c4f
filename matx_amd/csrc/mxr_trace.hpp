// Library log and trace ranges (implemented in trace.hip).
#pragma once

namespace mxr {
enum { LOG_TRACE = 0, LOG_DEBUG, LOG_INFO, LOG_WARN, LOG_ERROR, LOG_FATAL, LOG_OFF };

// True when a message at `level` would be written.
bool log_on(int level);
// printf-style message at `level` to the log destination, when enabled.
void log_msg(int level, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Opens a roctx range when `level` passes the trace level (true: pop it).
bool trace_push(const char *name, int level);
void trace_pop();
}  // namespace mxr
