// fvh_timing.hip -- per-kernel event timing behind fvh_timing_enable/read.
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fastvid_hip.h"
#include "fvh_timing.h"

namespace fvh {
namespace {

struct Pending {
  std::string name;
  hipEvent_t start, stop;
};

struct Totals {
  std::string name;
  int64_t launches;
  double total_ms;
};

std::mutex g_mutex;
bool g_enabled = false;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_free;
std::vector<Totals> g_totals;

hipEvent_t take_event() {
  if (!g_free.empty()) {
    hipEvent_t e = g_free.back();
    g_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

}  // namespace

bool timing_enabled() { return g_enabled; }

void timing_begin(const char* name, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(g_mutex);
  Pending p;
  p.name = name;
  p.start = take_event();
  p.stop = take_event();
  (void)hipEventRecord(p.start, stream);
  g_pending.push_back(p);
}

void timing_end(hipStream_t stream) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_pending.empty()) (void)hipEventRecord(g_pending.back().stop, stream);
}

}  // namespace fvh

using namespace fvh;

extern "C" {

int fvh_timing_enable(int32_t enable) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_enabled = enable != 0;
  return FVH_OK;
}

int fvh_timing_read(fvh_kernel_time* out, int32_t max, int32_t reset) {
  std::lock_guard<std::mutex> lock(g_mutex);
  for (Pending& p : g_pending) {
    (void)hipEventSynchronize(p.stop);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, p.start, p.stop);
    bool found = false;
    for (Totals& t : g_totals) {
      if (t.name == p.name) {
        t.launches += 1;
        t.total_ms += ms;
        found = true;
        break;
      }
    }
    if (!found) g_totals.push_back(Totals{p.name, 1, (double)ms});
    g_free.push_back(p.start);
    g_free.push_back(p.stop);
  }
  g_pending.clear();
  int n = 0;
  for (const Totals& t : g_totals) {
    if (n >= max) break;
    std::memset(out[n].name, 0, sizeof(out[n].name));
    std::strncpy(out[n].name, t.name.c_str(), sizeof(out[n].name) - 1);
    out[n].launches = t.launches;
    out[n].total_ms = t.total_ms;
    ++n;
  }
  if (reset) g_totals.clear();
  return n;
}

}  // extern "C"
