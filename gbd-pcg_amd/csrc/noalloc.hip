// The no-allocation solve: a stream-ordered forward solve whose device scratch
// (exchange records, result record, error word) is a fixed array in this
// library's device image, so a call allocates nothing.  For callers that cannot
// allocate device memory: code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc.  No counterpart in the reference.
//
// The array is per device (each GPU holds its own copy, at its own address, for
// the life of the process); its device address is looked up once per device with
// hipGetSymbolAddress and only that address is ever handed to kernels.  Every
// call on a device shares that device's array: calls that overlap on one device
// (different streams, or different host threads on unsynchronised streams) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

namespace gbdpcg {

void set_error(const char *fmt, ...);  // api.cpp

// 1 MiB: the exchange records of a single solve take 2 * G * rec_words(n) words,
// e.g. 4,352 at n = 14, N = 128 (16 workgroups, FP32) and about 72k at n = 28,
// N = 512 (FP64); systems needing more are rejected with GBDPCG_ERR_TOO_LARGE.
constexpr size_t kScratchWords = ((size_t)1 << 20) / sizeof(uint64_t);
static_assert(GBDPCG_NOALLOC_SCRATCH_BYTES == kScratchWords * sizeof(uint64_t), "scratch size");

static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
static std::mutex g_mu;
static uint64_t *g_ptr[kMaxDevices];  // device address of `scratch` on each device, once looked up

static int scratch_ptr(uint64_t **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("hipGetDevice failed: %s", hipGetErrorString(e));
    return GBDPCG_ERR_DEVICE;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d out of range for the no-allocation scratch", dev);
    return GBDPCG_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return GBDPCG_ERR_DEVICE;
    }
    g_ptr[dev] = static_cast<uint64_t *>(p);
  }
  *out = g_ptr[dev];
  return GBDPCG_OK;
}

static int noalloc_solve(int dtype, uint32_t n, uint32_t N, const void *S, const void *Pinv, const void *gamma,
                         void *lambda, const gbdpcg_config *cfg, gbdpcg_result *d_result) {
  uint64_t *p = nullptr;
  int rc = scratch_ptr(&p);
  if (rc < 0) return rc;
  return solve_on_scratch(dtype, n, N, S, Pinv, gamma, lambda, cfg, d_result, p, kScratchWords);
}

}  // namespace gbdpcg

extern "C" {

int gbdpcg_solve_device_noalloc_f32(uint32_t n, uint32_t N, const float *d_S, const float *d_Pinv,
                                    const float *d_gamma, float *d_lambda, const gbdpcg_config *cfg,
                                    gbdpcg_result *d_result) {
  return gbdpcg::noalloc_solve(GBDPCG_F32, n, N, d_S, d_Pinv, d_gamma, d_lambda, cfg, d_result);
}

int gbdpcg_solve_device_noalloc_f64(uint32_t n, uint32_t N, const double *d_S, const double *d_Pinv,
                                    const double *d_gamma, double *d_lambda, const gbdpcg_config *cfg,
                                    gbdpcg_result *d_result) {
  return gbdpcg::noalloc_solve(GBDPCG_F64, n, N, d_S, d_Pinv, d_gamma, d_lambda, cfg, d_result);
}

}  // extern "C"
