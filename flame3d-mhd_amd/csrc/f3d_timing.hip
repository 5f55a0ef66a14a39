// hipEvent timing of the library's own launches (f3d_timing_* in include/f3d.h).
#include <cstring>
#include <string>
#include <vector>

#include "f3d_internal.hpp"

namespace f3d {
namespace {
struct Rec {
  const char* name;
  hipEvent_t a, b;
  hipStream_t s;
};
bool g_on = false;
std::string g_only;        // when non-empty, only launches of this name are bracketed
std::vector<Rec> g_recs;   // recorded brackets since the last reset
std::vector<Rec> g_pool;   // event pairs available for reuse
size_t g_used = 0;         // brackets in use (prefix of g_recs)
}  // namespace

TimedScope::TimedScope(const char* name, hipStream_t s) : slot_(-1), s_(s) {
  if (!g_on) return;
  if (!g_only.empty() && g_only != name) return;
  if (g_used == g_recs.size()) {
    Rec r{name, nullptr, nullptr, s};
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
    g_recs.push_back(r);
  }
  Rec& r = g_recs[g_used];
  r.name = name;
  r.s = s;
  if (hipEventRecord(r.a, s) != hipSuccess) return;
  slot_ = (int)g_used++;
}

TimedScope::~TimedScope() {
  if (slot_ >= 0) (void)hipEventRecord(g_recs[slot_].b, s_);
}
}  // namespace f3d

using namespace f3d;
#define F3D_API extern "C" __attribute__((visibility("default")))

F3D_API int f3d_timing_enable(int32_t on) {
  g_on = on != 0;
  return F3D_OK;
}

F3D_API int f3d_timing_only(const char* name) {
  g_only = name ? name : "";
  return F3D_OK;
}

F3D_API int f3d_timing_reset(void) {
  g_used = 0;
  return F3D_OK;
}

F3D_API int f3d_timing_query(const char* name, double* total_ms, int64_t* launches) {
  if (!name || !total_ms || !launches) {
    set_error("timing_query: NULL argument");
    return F3D_E_NULL;
  }
  double tot = 0;
  int64_t n = 0;
  for (size_t i = 0; i < g_used; ++i) {
    const Rec& r = g_recs[i];
    if (std::strcmp(r.name, name) != 0) continue;
    F3D_HIP(hipEventSynchronize(r.b));
    float ms = 0;
    F3D_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    tot += ms;
    ++n;
  }
  *total_ms = tot;
  *launches = n;
  return F3D_OK;
}

F3D_API int f3d_timing_names(char* buf, size_t len) {
  if (!buf || len == 0) {
    set_error("timing_names: NULL buffer");
    return F3D_E_NULL;
  }
  std::string out;
  for (size_t i = 0; i < g_used; ++i) {
    const std::string nm = g_recs[i].name;
    bool seen = false;
    for (size_t j = 0; j < i && !seen; ++j) seen = nm == g_recs[j].name;
    if (seen) continue;
    if (!out.empty()) out += ",";
    out += nm;
  }
  std::strncpy(buf, out.c_str(), len - 1);
  buf[len - 1] = 0;
  return F3D_OK;
}
