// ls_noalloc.hip -- the Longstaff-Schwartz first pass for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  Its workspace is the library's own
// device array `scratch` below: one copy per GPU, allocated when the code
// object is loaded and held until the process ends.
//
// Every call on a GPU shares that GPU's array and the array keeps the last
// call's contents: the first pass writes every word it reads before reading
// it (running values, then the normal-equation partials date by date), and
// at the start of each one-launch pass a fill kernel arms every exchange slot
// with the "unset" pattern and another zeroes the flag word, so no state
// carries over.  Calls that overlap on one GPU (two streams, two host
// threads) race on the array and must not be issued.
#include <mutex>

#include "koo_common.h"

// Largest workspace the array serves: ls_workspace(paths) <= 16 MiB, i.e. up
// to 3 723 200 paths on a 256-CU device (the workspace's partials grow with
// the CU count).
constexpr size_t kLsScratchBytes = static_cast<size_t>(16) << 20;
static __device__ __align__(256) unsigned char scratch[kLsScratchBytes];

namespace koo {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on that device (with it current) and kept.  In host code the array's
// name is only a stand-in for the symbol; the address comes from the runtime.
int scratch_address(const char* who, void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return set_error(KOO_ERR_HIP, "%s: no current device (or device index >= %d)", who,
                     kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(KOO_ERR_HIP, "%s: hipGetSymbolAddress(scratch) failed on device %d: %s",
                       who, dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return KOO_OK;
}

}  // namespace
}  // namespace koo

using namespace koo;

extern "C" int ls_first_pass_noalloc_gpu(int n_ex, const float* basis_variables,
                                         const int* vars_each_step, int max_vars,
                                         const float* cash_flows, const float* exercise_values,
                                         const float* numeraires, int paths, int use_cross,
                                         int itm_only, const float* var_shift,
                                         const float* var_scale, double* coefficients,
                                         double* price_out, void* stream) {
  const char* who = "ls_first_pass_noalloc_gpu";
  size_t need = 0;
  if (int e = ls_workspace(paths, &need)) return e;
  if (need > kLsScratchBytes)
    return set_error(KOO_ERR_INVALID,
                     "%s: %d paths need a %zu-byte workspace; the library's scratch array holds "
                     "%zu bytes (use ls_first_pass_gpu with a workspace of your own)",
                     who, paths, need, kLsScratchBytes);
  void* ws = nullptr;
  if (int e = scratch_address(who, &ws)) return e;
  return ls_first_pass_gpu(n_ex, basis_variables, vars_each_step, max_vars, cash_flows,
                           exercise_values, numeraires, paths, use_cross, itm_only, var_shift,
                           var_scale, ws, coefficients, price_out, stream);
}
