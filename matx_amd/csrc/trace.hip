// Trace ranges and the library log.
//
// Trace ranges are the MI355X counterpart of MatX's NVTX ranges
// (include/matx/core/nvtx.h:48-168): named host-side ranges that a profiler
// shows beside the kernels (rocprofv3 --marker-trace).  They are emitted
// through roctx, which is loaded with dlopen the first time a range passes
// the level filter, so the library has no link-time dependency on it and
// costs nothing while no range is asked for.  Levels are MatX's: NONE 0,
// USER 1 (caller ranges), API 2, INTERNAL 3 (one range per kernel launch,
// named after the kernel), ALL.  The default level is API, as in MatX
// (nvtx.h:93); MXR_TRACE_LEVEL overrides it at start-up.
//
// The log is the counterpart of MatX's logger (include/matx/core/log.h:
// 110-401): levels TRACE 0 .. FATAL 5, OFF 6, from MXR_LOG_LEVEL (names or
// numbers), destination from MXR_LOG_DEST ("stderr", "stdout" or a file
// path, appended).  Unlike MatX it is OFF and on stderr by default: every
// failure is already returned as a status with a message, and stdout belongs
// to the caller.  Every failed call logs at ERROR, cache and plan events at
// DEBUG, kernel launches at TRACE.
#include <dlfcn.h>

#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <mutex>
#include <string>

#include "mxr_internal.hpp"
#include "mxr_trace.hpp"

namespace mxr {
namespace {

// ------------------------------------------------------------------ log
struct LogState {
  std::mutex mu;                    // guards dest and the writes
  std::atomic<int> level{LOG_OFF};  // read without the lock on every launch
  FILE *dest = stderr;
  bool own = false;  // dest is a file the library opened
};

int parse_log_level(const char *s, int dflt) {
  if (!s || !*s) return dflt;
  static const char *names[] = {"TRACE", "DEBUG", "INFO", "WARN", "ERROR", "FATAL", "OFF"};
  for (int i = 0; i <= LOG_OFF; ++i)
    if (strcasecmp(s, names[i]) == 0) return i;
  char *end = nullptr;
  const long v = strtol(s, &end, 10);
  if (end && *end == '\0' && v >= 0 && v <= LOG_OFF) return (int)v;
  return dflt;
}

// Opens `d` ("stderr", "stdout" or a path); false when the file cannot be
// opened (the destination is then left unchanged).
bool set_dest_locked(LogState &st, const char *d) {
  FILE *f = nullptr;
  bool own = false;
  if (!d || !*d || strcmp(d, "stderr") == 0) f = stderr;
  else if (strcmp(d, "stdout") == 0) f = stdout;
  else {
    f = fopen(d, "a");
    own = true;
  }
  if (!f) return false;
  if (st.own && st.dest) fclose(st.dest);
  st.dest = f;
  st.own = own;
  return true;
}

LogState &log_state() {
  static LogState *st = [] {
    auto *s = new LogState;  // never destroyed: logging may run during exit
    s->level = parse_log_level(getenv("MXR_LOG_LEVEL"), LOG_OFF);
    if (!set_dest_locked(*s, getenv("MXR_LOG_DEST"))) s->dest = stderr;
    return s;
  }();
  return *st;
}

const char *kLevelName[] = {"TRACE", "DEBUG", "INFO", "WARN", "ERROR", "FATAL", "OFF"};

// ---------------------------------------------------------------- trace
using PushFn = int (*)(const char *);
using PopFn = int (*)(void);
using StartFn = uint64_t (*)(const char *);
using StopFn = void (*)(uint64_t);

struct TraceState {
  std::mutex mu;
  std::atomic<int> level{MXR_TRACE_API};
  int loaded = 0;  // 0 not tried, 1 loaded, -1 unavailable
  PushFn push = nullptr;
  PopFn pop = nullptr;
  StartFn start = nullptr;
  StopFn stop = nullptr;
  std::map<int, uint64_t> live;  // caller id -> roctx range id
  int next_auto = 1 << 24;       // ids handed out for id < 0
};

TraceState &trace_state() {
  static TraceState *st = [] {
    auto *s = new TraceState;
    const char *e = getenv("MXR_TRACE_LEVEL");
    if (e && *e) {
      if (strcasecmp(e, "ALL") == 0) s->level = MXR_TRACE_ALL;
      else if (strcasecmp(e, "NONE") == 0) s->level = MXR_TRACE_NONE;
      else if (strcasecmp(e, "USER") == 0) s->level = MXR_TRACE_USER;
      else if (strcasecmp(e, "API") == 0) s->level = MXR_TRACE_API;
      else if (strcasecmp(e, "INTERNAL") == 0) s->level = MXR_TRACE_INTERNAL;
      else s->level = atoi(e);
    }
    return s;
  }();
  return *st;
}

// Loads roctx once; the caller holds st.mu.
bool trace_load_locked(TraceState &st) {
  if (st.loaded) return st.loaded > 0;
  const char *libs[] = {"librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so", "libroctx64.so.4",
                        "libroctx64.so"};
  for (const char *l : libs) {
    void *h = dlopen(l, RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    st.push = (PushFn)dlsym(h, "roctxRangePushA");
    st.pop = (PopFn)dlsym(h, "roctxRangePop");
    st.start = (StartFn)dlsym(h, "roctxRangeStartA");
    st.stop = (StopFn)dlsym(h, "roctxRangeStop");
    if (st.push && st.pop && st.start && st.stop) {
      st.loaded = 1;
      log_msg(LOG_DEBUG, "trace: roctx loaded from %s", l);
      return true;
    }
    dlclose(h);
  }
  st.loaded = -1;
  log_msg(LOG_WARN, "trace: no roctx library found; trace ranges are off");
  return false;
}
}  // namespace

bool log_on(int level) {
  const int cur = log_state().level.load(std::memory_order_relaxed);
  return level < LOG_OFF && level >= cur && cur < LOG_OFF;
}

void log_msg(int level, const char *fmt, ...) {
  if (!log_on(level)) return;
  char msg[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof(msg), fmt, ap);
  va_end(ap);
  const auto now = std::chrono::system_clock::now();
  const std::time_t t = std::chrono::system_clock::to_time_t(now);
  const int ms = (int)(std::chrono::duration_cast<std::chrono::milliseconds>(now.time_since_epoch()).count() % 1000);
  std::tm tm{};
  localtime_r(&t, &tm);
  char ts[32];
  std::strftime(ts, sizeof(ts), "%Y-%m-%dT%H:%M:%S", &tm);
  LogState &st = log_state();
  std::lock_guard<std::mutex> g(st.mu);
  const int lv = level < 0 ? 0 : (level > LOG_FATAL ? LOG_FATAL : level);
  fprintf(st.dest, "[%s.%03d] [mxr %s] %s\n", ts, ms, kLevelName[lv], msg);
  fflush(st.dest);
}

bool trace_push(const char *name, int level) {
  TraceState &st = trace_state();
  // the filter without the lock: every kernel launch passes through here
  if (level > st.level.load(std::memory_order_relaxed) || level <= MXR_TRACE_NONE) return false;
  std::lock_guard<std::mutex> g(st.mu);
  if (!trace_load_locked(st)) return false;
  st.push(name ? name : "");
  return true;
}

void trace_pop() {
  TraceState &st = trace_state();
  std::lock_guard<std::mutex> g(st.mu);
  if (st.loaded > 0) st.pop();
}
}  // namespace mxr

using namespace mxr;

extern "C" {

mxr_status mxr_log_set_level(int level) {
  MXR_REQUIRE(level >= LOG_TRACE && level <= LOG_OFF, MXR_INVALID_PARAMETER,
              "mxr_log_set_level: level %d outside [0, 6]", level);
  log_state().level.store(level);
  return MXR_SUCCESS;
}

int mxr_log_get_level(void) { return log_state().level.load(); }

mxr_status mxr_log_set_dest(const char *dest) {
  LogState &st = log_state();
  bool ok;
  {
    std::lock_guard<std::mutex> g(st.mu);
    ok = set_dest_locked(st, dest);
  }
  MXR_REQUIRE(ok, MXR_INVALID_PARAMETER, "mxr_log_set_dest: cannot open '%s'", dest ? dest : "");
  return MXR_SUCCESS;
}

void mxr_log_write(int level, const char *message) { log_msg(level, "%s", message ? message : ""); }

mxr_status mxr_trace_set_level(int level) {
  MXR_REQUIRE(level >= MXR_TRACE_NONE, MXR_INVALID_PARAMETER, "mxr_trace_set_level: level %d < 0", level);
  trace_state().level.store(level);
  return MXR_SUCCESS;
}

int mxr_trace_get_level(void) { return trace_state().level.load(); }

int mxr_trace_available(void) {
  TraceState &st = trace_state();
  std::lock_guard<std::mutex> g(st.mu);
  return trace_load_locked(st) ? 1 : 0;
}

int mxr_trace_push(const char *name, int level) { return trace_push(name, level) ? 1 : 0; }

void mxr_trace_pop(void) { trace_pop(); }

mxr_status mxr_trace_range_start(const char *name, int level, int *id) {
  MXR_REQUIRE(id != nullptr, MXR_INVALID_PARAMETER, "mxr_trace_range_start: null id");
  TraceState &st = trace_state();
  std::lock_guard<std::mutex> g(st.mu);
  if (*id < 0) *id = st.next_auto++;
  MXR_REQUIRE(st.live.find(*id) == st.live.end(), MXR_INVALID_PARAMETER,
              "mxr_trace_range_start: range id %d is already open", *id);
  if (level > st.level || level <= MXR_TRACE_NONE || !trace_load_locked(st)) return MXR_SUCCESS;
  st.live[*id] = st.start(name ? name : "");
  return MXR_SUCCESS;
}

mxr_status mxr_trace_range_end(int id) {
  TraceState &st = trace_state();
  std::lock_guard<std::mutex> g(st.mu);
  auto it = st.live.find(id);
  if (it == st.live.end()) return MXR_SUCCESS;  // filtered out when started, or never started
  st.stop(it->second);
  st.live.erase(it);
  return MXR_SUCCESS;
}

}  // extern "C"
