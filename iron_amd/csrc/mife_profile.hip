// Event-pair registry behind mife_profile_* (host code).
#include "mife_common.hpp"
#include "mife_profile.hpp"

#include <mutex>
#include <vector>

namespace mife {
namespace prof {

static const char* kNames[kNumKernels] = {"cg_spmv",          "cg_update",         "cg_init",
                                          "residual64",       "assemble_scalar",   "assemble_elastic",
                                          "gather_values",    "finite_elasticity", "grouped_spmv",
                                          "grouped_update",   "cg_spmv_f32",       "cg_update_f32",
                                          "assemble_scalar_f32", "assemble_elastic_f32",
                                          "minres_spmv",      "grouped_replace",   "assemble_batch",
                                          "finite_elasticity_residual"};

struct Interval {
  hipEvent_t a, b;
  int kernel;
  bool counted;
  bool closed;
};

static std::mutex g_mu;
static bool g_on = false;
static std::vector<Interval> g_open;
static std::vector<hipEvent_t> g_free;
static int g_count[kNumKernels];
static double g_ms[kNumKernels];

static hipEvent_t take_event() {
  if (!g_free.empty()) {
    hipEvent_t e = g_free.back();
    g_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

bool enabled() { return g_on; }

int begin(hipStream_t s) {
  if (!g_on) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t a = take_event(), b = take_event();
  if (!a || !b) return -1;
  if (hipEventRecord(a, s) != hipSuccess) return -1;
  g_open.push_back({a, b, -1, true, false});
  return (int)g_open.size() - 1;
}

void end(int ticket, hipStream_t s, int kernel) {
  if (ticket < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (ticket >= (int)g_open.size()) return;
  Interval& iv = g_open[ticket];
  iv.kernel = kernel;
  iv.closed = hipEventRecord(iv.b, s) == hipSuccess;
}

void discard(int ticket) {
  if (ticket < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (ticket < (int)g_open.size()) g_open[ticket].counted = false;
}

static int resolve() {
  for (auto& iv : g_open) {
    if (iv.closed && iv.counted && iv.kernel >= 0) {
      MIFE_CHECK_HIP(hipEventSynchronize(iv.b));
      float ms = 0.f;
      MIFE_CHECK_HIP(hipEventElapsedTime(&ms, iv.a, iv.b));
      g_count[iv.kernel] += 1;
      g_ms[iv.kernel] += ms;
    } else if (iv.closed) {
      hipEventSynchronize(iv.b);
    }
    g_free.push_back(iv.a);
    g_free.push_back(iv.b);
  }
  g_open.clear();
  return MIFE_OK;
}

}  // namespace prof
}  // namespace mife

using namespace mife;

extern "C" {

int mife_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(prof::g_mu);
  prof::g_on = on != 0;
  return MIFE_OK;
}

int mife_profile_reset(void) {
  std::lock_guard<std::mutex> lk(prof::g_mu);
  int rc = prof::resolve();
  for (int k = 0; k < prof::kNumKernels; ++k) {
    prof::g_count[k] = 0;
    prof::g_ms[k] = 0.0;
  }
  return rc;
}

int mife_profile_read(int n_max, char* names, int* counts, double* total_ms, int* n_out) {
  std::lock_guard<std::mutex> lk(prof::g_mu);
  int rc = prof::resolve();
  if (rc) return rc;
  const int n = n_max < prof::kNumKernels ? n_max : prof::kNumKernels;
  for (int k = 0; k < n; ++k) {
    if (names) {
      strncpy(names + 48 * k, prof::kNames[k], 47);
      names[48 * k + 47] = 0;
    }
    if (counts) counts[k] = prof::g_count[k];
    if (total_ms) total_ms[k] = prof::g_ms[k];
  }
  if (n_out) *n_out = n;
  return MIFE_OK;
}

}  // extern "C"
