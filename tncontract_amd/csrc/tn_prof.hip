// Launch timing for bench.py's roofline figure: HIP events around each
// instrumented launch, on the launch stream; summed per tag on collection.
#include "tn_prof.h"

#include <mutex>
#include <vector>

namespace tn {

namespace {
struct Rec {
  int tag;
  hipEvent_t a, b;
  double flops, bytes;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;
double g_extra[PROF_NTAGS][2] = {};

hipEvent_t get_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

bool prof_on() { return g_on; }

int prof_start(int tag, hipStream_t st, double flops, double bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_on) return -1;
  Rec r{tag, get_event(), get_event(), flops, bytes};
  if (!r.a || !r.b) return -1;
  if (hipEventRecord(r.a, st) != hipSuccess) return -1;
  g_recs.push_back(r);
  return (int)g_recs.size() - 1;
}

void prof_stop(int slot, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (slot < 0 || slot >= (int)g_recs.size()) return;
  (void)hipEventRecord(g_recs[slot].b, st);
}

void prof_add(int tag, double flops, double bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_on || tag < 0 || tag >= PROF_NTAGS) return;
  g_extra[tag][0] += flops;
  g_extra[tag][1] += bytes;
}

}  // namespace tn

using namespace tn;

extern "C" int tn_prof_enable(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_on = on != 0;
  return TN_OK;
}

extern "C" int tn_prof_collect(int tag, double* total_ms, int64_t* launches, double* flops,
                               double* bytes, int reset) {
  TN_REQUIRE(tag >= 0 && tag < PROF_NTAGS, "tn_prof_collect: bad tag %d", tag);
  std::lock_guard<std::mutex> lk(g_mu);
  double ms = 0.0, fl = 0.0, by = 0.0;
  int64_t n = 0;
  for (auto& r : g_recs) {
    if (r.tag != tag) continue;
    TN_HIP(hipEventSynchronize(r.b));
    float t = 0.f;
    TN_HIP(hipEventElapsedTime(&t, r.a, r.b));
    ms += t;
    fl += r.flops;
    by += r.bytes;
    ++n;
  }
  fl += g_extra[tag][0];
  by += g_extra[tag][1];
  if (total_ms) *total_ms = ms;
  if (launches) *launches = n;
  if (flops) *flops = fl;
  if (bytes) *bytes = by;
  if (reset) {
    for (auto& r : g_recs) {
      (void)hipEventSynchronize(r.b);
      g_pool.push_back(r.a);
      g_pool.push_back(r.b);
    }
    g_recs.clear();
    for (auto& e : g_extra) e[0] = e[1] = 0.0;
  }
  return TN_OK;
}
