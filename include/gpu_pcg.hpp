// gpu_pcg.hpp — C++ drop-in facade over the C ABI (gbdpcg.h), mirroring the
// reference's host interface so existing callers keep compiling:
//
//   reference (file:line)                                   here
//   pcg_config<T>            include/types.cuh:18-35        gbdpcg::pcg_config<T> (same field names/defaults)
//   csr_t<T>                 include/types.cuh:7-15         gbdpcg::csr_t<T>
//   solvePCG(csr_t*, ...)    include/interface.cuh:8-20     solvePCG(csr_t*, ...)  (implemented; the reference exits 12)
//   solvePCG(h_S, h_Pinv, h_gamma, h_lambda, n, N, cfg)
//                            include/interface.cuh:24-91    solvePCG(...)  host pointers
//   solvePCG(n, N, d_S, d_Pinv, d_gamma, d_lambda, d_r, d_p, d_v_temp, d_eta_new_temp, cfg)
//                            include/interface.cuh:94-146   solvePCG(...)  device pointers (scratch unused)
//   pcg_solve(cbtd_t*, ...)  README.md:41-54 (documented,   pcg_solve(cbtd_t<T>*, ...)
//                            not implemented there)
//   pcgSharedMemSize<T>      include/pcg.cuh:13-20          pcgSharedMemSize<T>
//   checkPcgOccupancy<T>     include/pcg.cuh:23-49          checkPcgOccupancy<T>
//
// Behaviour that differs on purpose: errors come back as negative return values
// (GBDPCG_ERR_*; gbdpcg_last_error() has the message) instead of exit(); the
// host solve returns the iteration count (the reference returns 1,
// interface.cuh:90); stateSize and knotPoints are runtime values (no
// -DSTATE_SIZE/-DKNOT_POINTS, interface.cuh:112); pcg_grid.x, when not the
// reference default, is the number of workgroups per system.
#pragma once

#include <stdint.h>

#include "gbdpcg.h"

namespace gbdpcg {

// CUDA's dim3 without the CUDA headers (the reference config carries two).
struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1) : x(x_), y(y_), z(z_) {}
};

namespace pcg_constants {  // include/constants.cuh:13-20 (inline: no ODR clash across TUs)
inline constexpr uint32_t DEFAULT_MAX_PCG_ITER = 25;
template <typename T>
inline constexpr T DEFAULT_EPSILON = T(1e-6);
inline constexpr dim3 DEFAULT_GRID{128};
inline constexpr dim3 DEFAULT_BLOCK{64};
}  // namespace pcg_constants

template <typename T>
struct csr_t {
  uint32_t *row_ptr;
  uint32_t *col_ind;
  T *val;
  uint32_t rows;
  uint32_t cols;
  uint32_t nnz;
};

template <typename T>
struct pcg_config {
  T pcg_exit_tol;
  uint32_t pcg_max_iter;
  dim3 pcg_grid;
  dim3 pcg_block;
  int empty_pinv;
  // extensions (defaults keep the reference behaviour)
  int exit_mode = GBDPCG_EXIT_ABS_ETA;
  int precond = GBDPCG_PRECOND_SS;
  int recurrence = GBDPCG_RECUR_REFERENCE;
  void *stream = nullptr;
  uint32_t kernel = GBDPCG_KERNEL_AUTO;
  uint32_t knots_per_wg = 0;    // 0 = automatic (tuned by gbdpcg_tune_single, else the planner)
  uint32_t threads = 0;         // register kernels: 64 (row broadcast), 256 or 512; 0 = automatic
  uint32_t exchange_scope = GBDPCG_EXCHANGE_AUTO;  // GBDPCG_EXCHANGE_DEVICE: device-scope stores always
  gbdpcg_result last_result{};  // filled by the device-pointer solve

  pcg_config(T exit_tol = pcg_constants::DEFAULT_EPSILON<T>,
             uint32_t max_iter = pcg_constants::DEFAULT_MAX_PCG_ITER, dim3 grid = pcg_constants::DEFAULT_GRID,
             dim3 block = pcg_constants::DEFAULT_BLOCK, int empty_pinv_ = 1)
      : pcg_exit_tol(exit_tol), pcg_max_iter(max_iter), pcg_grid(grid), pcg_block(block), empty_pinv(empty_pinv_) {}
};

// Compressed block-tridiagonal matrix pair: S and (optionally) Pinv, host or
// device pointers depending on the call.  README.md:42 names cbtd_t without
// defining it; Pinv == nullptr means "build the preconditioner from S".
template <typename T>
struct cbtd_t {
  T *S;
  T *Pinv;
};

namespace detail {
template <typename T>
inline gbdpcg_config to_c(const pcg_config<T> *cfg) {
  gbdpcg_config c;
  gbdpcg_config_init(&c);
  if (cfg) {
    c.exit_tol = (double)cfg->pcg_exit_tol;
    c.max_iter = cfg->pcg_max_iter;
    // the reference ignores pcg_grid (interface.cuh:134); any other value picks
    // the number of workgroups per system
    c.grid[0] = cfg->pcg_grid.x == pcg_constants::DEFAULT_GRID.x ? 0u : cfg->pcg_grid.x;
    c.block[0] = cfg->pcg_block.x;
    c.empty_pinv = cfg->empty_pinv;
    c.exit_mode = cfg->exit_mode;
    c.precond = cfg->precond;
    c.recurrence = cfg->recurrence;
    c.stream = cfg->stream;
    c.kernel = cfg->kernel;
    c.knots_per_wg = cfg->knots_per_wg;
    c.threads = cfg->threads;
    c.exchange_scope = cfg->exchange_scope;
  }
  return c;
}
template <typename T>
struct dt;
template <>
struct dt<float> {
  static constexpr int code = GBDPCG_F32;
};
template <>
struct dt<double> {
  static constexpr int code = GBDPCG_F64;
};
}  // namespace detail

// Host pointers, called as the reference example does (solvePCG<float>(h_S, ...),
// examples/pcg_solve.cu:35-41).  Returns the iteration count, or a negative
// GBDPCG_ERR_*.
template <typename T>
inline int solvePCG(T *h_S, T *h_Pinv, T *h_gamma, T *h_lambda, unsigned stateSize, unsigned knotPoints,
                    pcg_config<T> *config) {
  const gbdpcg_config c = detail::to_c(config);
  if constexpr (detail::dt<T>::code == GBDPCG_F64)
    return gbdpcg_solve_f64(h_S, h_Pinv, h_gamma, h_lambda, stateSize, knotPoints, &c);
  else
    return gbdpcg_solve_f32(h_S, h_Pinv, h_gamma, h_lambda, stateSize, knotPoints, &c);
}

// Device pointers (the reference's scratch vectors are accepted and not used).
// Waits for config->stream; the full result record lands in config->last_result.
template <typename T>
inline int solvePCG(uint32_t state_size, uint32_t knot_points, T *d_S, T *d_Pinv, T *d_gamma, T *d_lambda,
                    T * /*d_r*/, T * /*d_p*/, T * /*d_v_temp*/, T * /*d_eta_new_temp*/, pcg_config<T> *config) {
  const gbdpcg_config c = detail::to_c(config);
  gbdpcg_result r{};
  int rc;
  if constexpr (detail::dt<T>::code == GBDPCG_F64)
    rc = gbdpcg_solve_device_f64(state_size, knot_points, d_S, d_Pinv, d_gamma, d_lambda, &c, &r);
  else
    rc = gbdpcg_solve_device_f32(state_size, knot_points, d_S, d_Pinv, d_gamma, d_lambda, &c, &r);
  if (config) config->last_result = r;
  return rc;
}

// CSR input (host pointers, 0-based, nN x nN).  h_Pinv may be nullptr.
template <typename T>
inline int solvePCG(csr_t<T> *h_S, csr_t<T> *h_Pinv, T *h_gamma, T *h_lambda, unsigned stateSize,
                    unsigned knotPoints, pcg_config<T> *config) {
  if (!h_S) return GBDPCG_ERR_INVALID;
  const gbdpcg_config c = detail::to_c(config);
  const uint32_t *pr = h_Pinv ? h_Pinv->row_ptr : nullptr;
  const uint32_t *pc = h_Pinv ? h_Pinv->col_ind : nullptr;
  const T *pv = h_Pinv ? h_Pinv->val : nullptr;
  if constexpr (detail::dt<T>::code == GBDPCG_F64)
    return gbdpcg_solve_csr_f64(h_S->row_ptr, h_S->col_ind, h_S->val, pr, pc, pv, h_gamma, h_lambda, stateSize,
                                knotPoints, &c);
  else
    return gbdpcg_solve_csr_f32(h_S->row_ptr, h_S->col_ind, h_S->val, pr, pc, pv, h_gamma, h_lambda, stateSize,
                                knotPoints, &c);
}

// The documented entry point (README.md:41-54): host pointers, cbtd_t holds S and
// optionally Pinv; with Pinv == nullptr the config's preconditioner is built on the
// device.  Returns the iteration count or a negative GBDPCG_ERR_*.
template <typename T>
inline int pcg_solve(cbtd_t<T> *h_S, T *h_gamma, T *h_lambda, unsigned stateSize, unsigned knotPoints,
                     pcg_config<T> *config = nullptr) {
  if (!h_S) return GBDPCG_ERR_INVALID;
  pcg_config<T> dflt;
  return solvePCG(h_S->S, h_S->Pinv, h_gamma, h_lambda, stateSize, knotPoints, config ? config : &dflt);
}

// LDS bytes per workgroup of the default layout (up to 8 knots per workgroup).
template <typename T>
inline size_t pcgSharedMemSize(uint32_t state_size, uint32_t knot_points) {
  const uint32_t k = knot_points < 8 ? knot_points : 8;
  const uint32_t g = k ? (knot_points + k - 1) / k : 0;
  return gbdpcg_lds_bytes(detail::dt<T>::code, state_size, k, g);
}

// true when the system fits the device (LDS and co-resident workgroups); never exits.
template <typename T>
inline bool checkPcgOccupancy(uint32_t state_size, uint32_t knot_points) {
  uint32_t max_wgs = 0;
  return gbdpcg_check_occupancy(detail::dt<T>::code, state_size, knot_points, &max_wgs) == GBDPCG_OK;
}

}  // namespace gbdpcg
