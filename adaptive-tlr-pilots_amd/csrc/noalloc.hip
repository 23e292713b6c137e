// The allocation-free TLR log-likelihood: tlr_loglik without a context, its device workspace a file-scope
// static array of TLR_NOALLOC_BYTES per GPU (held until the process ends).
//
// The array's name is only a host-side stand-in for the device object; its device address comes from
// hipGetSymbolAddress, once per device with that device current (on the entry point's first call there),
// and only that address is handed on (carved into the evaluation's buffers, which the kernels take as
// arguments).  Every call on one GPU works in that GPU's one array: the evaluation clears what has to start
// at zero (status words, list counters) with fill kernels, and calls on one GPU from this process are
// serialised by a host mutex.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/tlrgp.h"
#include "tlr_kernels.hpp"

static __device__ __align__(256) unsigned char scratch[TLR_NOALLOC_BYTES];

namespace {
constexpr int MAX_DEV = 64;
constexpr int GRID_CAP = 4;   // persistent recompression grids (the problems that fit are small)
std::mutex g_mu[MAX_DEV];
void* g_scratch[MAX_DEV] = {};
}  // namespace

extern "C" int tlr_loglik_noalloc(const double* S, const double* y, int64_t n, int32_t b, const double theta[3],
                                  int32_t rule, double value, int32_t precision, double out3[3],
                                  int32_t* ranks_sigma, int32_t* ranks_factor, int32_t* failed_at) {
  if (n > TLR_NOALLOC_MAX_N || b > TLR_NOALLOC_MAX_B) {
    const std::string m = "tlr_loglik_noalloc: N=" + std::to_string(n) + ", b=" + std::to_string(b) +
                          " exceed its static workspace (N <= " + std::to_string(TLR_NOALLOC_MAX_N) +
                          ", b <= " + std::to_string(TLR_NOALLOC_MAX_B) + ")";
    return tlr_set_error(TLR_ERR_ARG, m.c_str());
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEV)
    return tlr_set_error(TLR_ERR_NODEVICE, "tlr_loglik_noalloc: no current HIP device");
  std::lock_guard<std::mutex> lock(g_mu[dev]);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      const std::string m = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return tlr_set_error(TLR_ERR_HIP, m.c_str());
    }
    g_scratch[dev] = p;
  }
  return tlr_loglik_in_arena(g_scratch[dev], TLR_NOALLOC_BYTES, GRID_CAP, S, y, n, b, theta, rule, value, precision,
                             out3, ranks_sigma, ranks_factor, failed_at);
}
