// Allocation-free variant of the one-launch forest step, for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops).
// It has no counterpart in the reference.
#include <mutex>

#include "smk_internal.h"

// The work queue of smk_forest_step in a file-scope device array instead of a
// caller workspace.  Each GPU holds its own copy until the process ends; its
// address on a device is taken once with hipGetSymbolAddress, with that
// device current, and kept per device: host code passes that address, never
// the array's name, to kernels.  Calls on one GPU share the array (each call
// clears what it uses first), so calls that overlap on one GPU race.
static __device__ __align__(256) int32_t scratch[64];
constexpr size_t SMK_STEP_SCRATCH_BYTES = sizeof(int32_t) * 64;

namespace {
constexpr int SMK_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_ptr[SMK_MAX_DEVICES] = {};

int step_scratch(int32_t** out) {
  int dev = 0;
  SMK_HIP_CHECK(hipGetDevice(&dev));
  SMK_REQUIRE(dev >= 0 && dev < SMK_MAX_DEVICES, SMK_ERR_ARG, "device %d beyond %d", dev, SMK_MAX_DEVICES);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    SMK_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    SMK_REQUIRE(ptr != nullptr, SMK_ERR_HIP, "no device address for the scratch array");
    g_scratch_ptr[dev] = (int32_t*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return SMK_OK;
}
}  // namespace

extern "C" size_t smk_forest_step_noalloc_scratch_bytes(void) { return SMK_STEP_SCRATCH_BYTES; }

extern "C" int smk_forest_step_noalloc(smk_forest_plan* plan, const smk_skewers* sk, const smk_box_view* box,
                                       const float* filt, const int64_t* filt_off, int64_t set_stride,
                                       int32_t n_sets, const int32_t* filt_set, const float* scale, uint64_t seed,
                                       const float* abcg, int64_t abcg_rows, int32_t max_npix, float* flux,
                                       float* delta_l, float* eta_par, float* delta_s, int32_t* status,
                                       void* stream) {
  SMK_REQUIRE(smk_forest_step_workspace_bytes() <= SMK_STEP_SCRATCH_BYTES, SMK_ERR_RANGE,
              "the step needs %zu bytes of scratch, the static array holds %zu",
              smk_forest_step_workspace_bytes(), SMK_STEP_SCRATCH_BYTES);
  int32_t* ws = nullptr;
  const int rc = step_scratch(&ws);
  if (rc != SMK_OK) return rc;
  return smk_forest_step(plan, sk, box, filt, filt_off, set_stride, n_sets, filt_set, scale, seed, abcg, abcg_rows,
                         max_npix, flux, delta_l, eta_par, delta_s, status, ws, stream);
}
