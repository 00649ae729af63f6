/*
 * gbdpcg.h — C ABI of the MI355X (gfx950) block-tridiagonal PCG solver.
 *
 * Plain C: pointers, sizes and POD structs only.  Every entry point returns an
 * int: a non-negative value is a success (for the solve calls: the iteration
 * count of the forward solve), a negative value is one of GBDPCG_ERR_*.
 * Nothing in this library calls exit() or abort().
 *
 * Storage ("cbtd", compressed block-tridiagonal), as consumed by the reference
 * kernel (include/pcg.cuh:103-110, include/utils.cuh:46-85):
 *   a matrix with N block rows of n x n blocks is stored as N * 3*n*n values;
 *   block row k occupies [k*3n^2, (k+1)*3n^2) and holds, in order, the blocks
 *   (k,k-1), (k,k), (k,k+1), each column-major.  Equivalently block row k is an
 *   n x 3n column-major matrix: element (r, c) sits at k*3n^2 + c*n + r.
 *   The (0,-1) and (N-1,N) blocks are padding: never read, written as zeros
 *   where this library produces cbtd output.
 *   Vectors (gamma, lambda, ...) are knot-major: x_k occupies [k*n, (k+1)*n).
 *
 * Reference interface map (file:line in the reference tree):
 *   gbdpcg_solve_f32/_f64              <- solvePCG(h_S, h_Pinv, h_gamma, h_lambda, n, N, cfg)
 *                                         include/interface.cuh:24-91
 *   gbdpcg_solve_device_f32/_f64       <- solvePCG(n, N, d_S, d_Pinv, d_gamma, d_lambda,
 *                                         d_r, d_p, d_v_temp, d_eta_new_temp, cfg)
 *                                         include/interface.cuh:94-146
 *   gbdpcg_solve_csr_f32/_f64          <- solvePCG(csr_t*, csr_t*, ...) include/interface.cuh:8-20
 *   gbdpcg_config                      <- pcg_config<T> include/types.cuh:18-35,
 *                                         defaults include/constants.cuh:14-20
 *   gbdpcg_lds_bytes                   <- pcgSharedMemSize<T> include/pcg.cuh:13-20
 *   gbdpcg_check_occupancy             <- checkPcgOccupancy<T> include/pcg.cuh:23-49
 * Everything else (async/device-result solve, batched plans, adjoint, on-device
 * preconditioner formation, problem generation) extends the reference.
 */
#ifndef GBDPCG_H
#define GBDPCG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes (negative) ---- */
#define GBDPCG_OK 0
#define GBDPCG_ERR_INVALID (-1)     /* null pointer, zero size, bad enum ...        */
#define GBDPCG_ERR_TOO_LARGE (-2)   /* does not fit LDS / co-resident workgroups     */
#define GBDPCG_ERR_DEVICE (-3)      /* a HIP runtime call failed (or no device)       */
#define GBDPCG_ERR_TIMEOUT (-4)     /* a device-side exchange wait hit its bound      */
#define GBDPCG_ERR_NOMEM (-5)       /* device or host allocation failed               */
#define GBDPCG_ERR_UNSUPPORTED (-6) /* input form not supported (see the CSR entry)   */

/* ---- per-system solver status (gbdpcg_result.status) ---- */
#define GBDPCG_CONVERGED 0 /* exit test met                                     */
#define GBDPCG_MAX_ITER 1  /* max_iter iterations without meeting the exit test  */
#define GBDPCG_BREAKDOWN 2 /* p'Sp == 0 or r'z == 0 with r != 0                  */
#define GBDPCG_NONFINITE 3 /* NaN/Inf appeared in a reduction                    */
#define GBDPCG_TIMED_OUT 4 /* a cross-workgroup wait exceeded its bound           */

/* ---- enums ---- */
#define GBDPCG_F32 0
#define GBDPCG_F64 1

/* exit test: ABS_ETA is the reference's |r'z| < tol (include/pcg.cuh:195);
 * REL_RESID is ||r||_2 <= tol * ||rhs||_2 (evaluated as r'r <= tol^2 rhs'rhs). */
#define GBDPCG_EXIT_ABS_ETA 0
#define GBDPCG_EXIT_REL_RESID 1

/* preconditioner used when no Pinv is passed (Pinv == NULL or empty_pinv != 0
 * on the entry points that document it). */
#define GBDPCG_PRECOND_JACOBI 1 /* blockdiag(theta_k^-1)                               */
#define GBDPCG_PRECOND_SS 2     /* symmetric stair: theta_k^-1 on the diagonal,
                                   -theta_k^-1 phi_k theta_{k-1}^-1 below it,
                                   its transpose above                                  */

/* recurrence: REFERENCE is the reference iteration (include/pcg.cuh:118-208), two
 * grid-wide reductions per iteration.  FUSED is the Chronopoulos-Gear form of the
 * same method, one reduction per iteration (identical in exact arithmetic).  FUSED
 * needs at least 2 knots per workgroup: a system that no layout gives that (N = 1,
 * or LDS-resident rows too large for a CU's LDS at 2 or 3 knots per workgroup, as
 * FP64 at odd N from n = 25 on and at any N > 2 for n = 32) runs the reference
 * iteration instead, in single solves and batch plans alike. */
#define GBDPCG_RECUR_REFERENCE 0
#define GBDPCG_RECUR_FUSED 1

typedef struct gbdpcg_config {
  double exit_tol;    /* pcg_exit_tol, default 1e-6 (constants.cuh:16)                  */
  uint32_t max_iter;  /* pcg_max_iter, default 25 (constants.cuh:15)                    */
  uint32_t grid[3];   /* pcg_grid: grid[0] = workgroups per system, 0 = automatic        */
  uint32_t block[3];  /* pcg_block: accepted for layout parity; the kernel uses 256      */
  int32_t empty_pinv; /* 1 = Pinv argument absent: build it on device (see precond)     */
  int32_t exit_mode;  /* GBDPCG_EXIT_*, default ABS_ETA (the reference test)             */
  int32_t precond;    /* GBDPCG_PRECOND_* used when Pinv is absent, default SS           */
  int32_t recurrence; /* GBDPCG_RECUR_*, default REFERENCE                               */
  void *stream;       /* hipStream_t; NULL = the default stream                          */
  uint32_t kernel;    /* GBDPCG_KERNEL_*: 0 = automatic                                  */
  uint32_t knots_per_wg; /* cap on knots per workgroup, 0 = automatic (tuned or planner) */
  uint32_t threads;   /* threads per workgroup of the register kernels: 64 (row broadcast
                         only: one wave per rank), 256 or 512; 0 = automatic          */
  uint32_t batch_split; /* batch plans: 0 = one persistent launch for every system;
                           1 = one launch per kernel class present (rows of n <= 14,
                           rows of n = 28, LDS-resident rows), back to back          */
  uint32_t exchange_scope; /* GBDPCG_EXCHANGE_*: how ranks publish their exchange words */
  uint32_t reserved[3];
} gbdpcg_config;

/* exchange stores (gbdpcg_config.exchange_scope).  AUTO: workgroup-scope stores,
 * which complete in the shared L2 of one XCD, when the first exchange of a solve
 * finds every rank of the system on one XCD, else device-scope stores.  DEVICE:
 * device-scope stores always (written through to memory; slower, and independent
 * of the placement).  The environment variable GBDPCG_DEVICE_SCOPE_EXCHANGE=1
 * forces DEVICE for every call of the process. */
#define GBDPCG_EXCHANGE_AUTO 0
#define GBDPCG_EXCHANGE_DEVICE 1

/* kernel selection (gbdpcg_config.kernel) */
#define GBDPCG_KERNEL_AUTO 0
#define GBDPCG_KERNEL_LDS 1      /* LDS-resident block rows, both recurrences, any n  */
#define GBDPCG_KERNEL_REGISTER 2 /* register-resident rows (column chunks),
                                    n in {2, 6, 12, 14, 28}                           */
#define GBDPCG_KERNEL_ROWCAST 3  /* register-resident rows, one lane per row and 16
                                    lanes per knot, the vector window broadcast by
                                    DPP within each 16-lane row; n in {12, 14}        */

typedef struct gbdpcg_result {
  int32_t iters;      /* forward iterations (reference counting, pcg.cuh:195,212)   */
  int32_t status;     /* GBDPCG_CONVERGED ...                                        */
  int32_t adj_iters;  /* adjoint iterations (0 if no adjoint solve)                  */
  int32_t adj_status;
  double eta;         /* final r'z of the forward solve                              */
  double rr;          /* final r'r of the forward solve                              */
  double adj_eta;
  double adj_rr;
} gbdpcg_result;

/* Fills cfg with the reference defaults (tol 1e-6, max_iter 25, grid 128, block 64,
 * empty_pinv 1, ABS_ETA, SS, REFERENCE, default stream). */
void gbdpcg_config_init(gbdpcg_config *cfg);

/* Library version string. */
const char *gbdpcg_version(void);

/* Dynamic LDS bytes of one workgroup owning `knots_per_wg` knots of a stateSize-n
 * system with `wgs` workgroups per system (the reference's pcgSharedMemSize,
 * include/pcg.cuh:13-20, computed from the actual carve-up).  0 on bad input. */
size_t gbdpcg_lds_bytes(int dtype, uint32_t n, uint32_t knots_per_wg, uint32_t wgs);

/* Latency floor of one exchange round (extension, diagnostic): `wgs` workgroups
 * of `threads` threads (64, 256 or 512) run `rounds` (>= 20) rounds of the
 * solvers' exchange alone -- every rank publishes three words, waits for every
 * rank's words, then a workgroup barrier -- placed as the solvers place a system
 * (one XCD when it fits).  *us_per_round gets the time per round from the
 * difference of a short and a long run; *xcd_local (may be NULL) whether the
 * rounds used XCD-local stores.  What one PCG iteration cannot go below. */
int gbdpcg_exchange_floor(uint32_t wgs, uint32_t threads, uint32_t rounds, double *us_per_round,
                          uint32_t *xcd_local);

/* Measured resource rates (extension, diagnostic; no counterpart in the
 * reference): the rates the solvers' rooflines are read against, each from a
 * streaming kernel over the whole device -- *valu_f32_tflops: FP32 FMA
 * (one lane-op per FMA, no packed-FP32 instructions); *lds_read_gbs: 16-byte
 * conflict-free LDS reads; *hbm_read_gbs: one pass over a 2 GiB buffer.
 * Any pointer may be NULL (that rate is not measured).  Allocates and frees
 * its own buffers; synchronous. */
int gbdpcg_measure_rates(double *valu_f32_tflops, double *lds_read_gbs, double *hbm_read_gbs);

/* Co-residency check (the reference's checkPcgOccupancy, include/pcg.cuh:23-49, which
 * exits the process; this returns a code).  On success *max_wgs receives the number
 * of workgroups per system that can be co-resident for this (n, N).  */
int gbdpcg_check_occupancy(int dtype, uint32_t n, uint32_t N, uint32_t *max_wgs);

/* ---------------------------------------------------------------------------
 * Drop-in solves
 * ------------------------------------------------------------------------- */

/* Host-pointer solve: replaces solvePCG(h_S, h_Pinv, h_gamma, h_lambda, n, N, cfg),
 * include/interface.cuh:24-91.  Copies in, solves on device 0 (current device),
 * copies lambda back.  h_S, h_Pinv, h_gamma are read only (the reference writes
 * S and gamma back unchanged, interface.cuh:77,79).  h_Pinv may be NULL: the
 * preconditioner cfg->precond is then built on the device from S.
 * Returns the iteration count (the reference wrapper returns 1, interface.cuh:90). */
int gbdpcg_solve_f32(const float *h_S, const float *h_Pinv, const float *h_gamma, float *h_lambda,
                     uint32_t stateSize, uint32_t knotPoints, const gbdpcg_config *cfg);
int gbdpcg_solve_f64(const double *h_S, const double *h_Pinv, const double *h_gamma,
                     double *h_lambda, uint32_t stateSize, uint32_t knotPoints,
                     const gbdpcg_config *cfg);

/* Device-pointer solve: replaces solvePCG(n, N, d_S, d_Pinv, d_gamma, d_lambda, d_r,
 * d_p, d_v_temp, d_eta_new_temp, cfg), include/interface.cuh:94-146.  The
 * reference's scratch vectors are not needed (the iteration vectors stay on chip).
 * Launches on cfg->stream, waits for it, returns the iteration count; if `result`
 * is not NULL it receives the full result record.  d_Pinv may be NULL (see above). */
int gbdpcg_solve_device_f32(uint32_t stateSize, uint32_t knotPoints, const float *d_S,
                            const float *d_Pinv, const float *d_gamma, float *d_lambda,
                            const gbdpcg_config *cfg, gbdpcg_result *result);
int gbdpcg_solve_device_f64(uint32_t stateSize, uint32_t knotPoints, const double *d_S,
                            const double *d_Pinv, const double *d_gamma, double *d_lambda,
                            const gbdpcg_config *cfg, gbdpcg_result *result);

/* Asynchronous device-pointer solve: as above but does not wait; the result record
 * is written to device memory d_result (may be NULL) in stream order.  Returns 0 on a
 * successful launch.  Safe to capture in a hipGraph (no allocation or memset on this
 * path once a workspace exists for the stream; call it once outside capture first). */
int gbdpcg_solve_device_async_f32(uint32_t stateSize, uint32_t knotPoints, const float *d_S,
                                  const float *d_Pinv, const float *d_gamma, float *d_lambda,
                                  const gbdpcg_config *cfg, gbdpcg_result *d_result);
int gbdpcg_solve_device_async_f64(uint32_t stateSize, uint32_t knotPoints, const double *d_S,
                                  const double *d_Pinv, const double *d_gamma, double *d_lambda,
                                  const gbdpcg_config *cfg, gbdpcg_result *d_result);

/* No-allocation solve (extension; no counterpart in the reference).  For callers
 * that cannot allocate device memory -- code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc -- the same solve as
 * gbdpcg_solve_device_async_*, with its device scratch (exchange records, result
 * record, error word) in a fixed array of GBDPCG_NOALLOC_SCRATCH_BYTES held by
 * the library on each device for the life of the process: the call allocates
 * nothing and records no memset (a fill kernel clears the scratch).  d_Pinv is
 * required (forming it would need device memory).  The result record goes to
 * d_result when it is not NULL (status GBDPCG_TIMED_OUT there reports a failed
 * exchange wait).  Systems whose exchange records exceed the scratch are
 * rejected with GBDPCG_ERR_TOO_LARGE before any launch (n = 14 FP32 up to about
 * N = 3800; n = 28 FP64 up to about N = 900).  Every call on one device shares
 * that device's scratch: calls that can overlap on one device race, so keep them
 * on one stream (or synchronise between them).  Returns 0 after the launch. */
#define GBDPCG_NOALLOC_SCRATCH_BYTES (1u << 20)
int gbdpcg_solve_device_noalloc_f32(uint32_t stateSize, uint32_t knotPoints, const float *d_S,
                                    const float *d_Pinv, const float *d_gamma, float *d_lambda,
                                    const gbdpcg_config *cfg, gbdpcg_result *d_result);
int gbdpcg_solve_device_noalloc_f64(uint32_t stateSize, uint32_t knotPoints, const double *d_S,
                                    const double *d_Pinv, const double *d_gamma, double *d_lambda,
                                    const gbdpcg_config *cfg, gbdpcg_result *d_result);

/* CSR input: the reference stubs this overload (include/interface.cuh:8-20 prints
 * "NOT IMPLEMENTED" and exits with 12).  Here the CSR matrices (host pointers,
 * 0-based, square nN x nN) are converted to cbtd on the host and solved; entries
 * outside the block-tridiagonal pattern give GBDPCG_ERR_UNSUPPORTED.  h_Pinv_* may
 * be NULL (then cfg->precond is built). */
int gbdpcg_solve_csr_f32(const uint32_t *S_row_ptr, const uint32_t *S_col_ind, const float *S_val,
                         const uint32_t *P_row_ptr, const uint32_t *P_col_ind, const float *P_val,
                         const float *h_gamma, float *h_lambda, uint32_t stateSize,
                         uint32_t knotPoints, const gbdpcg_config *cfg);
int gbdpcg_solve_csr_f64(const uint32_t *S_row_ptr, const uint32_t *S_col_ind, const double *S_val,
                         const uint32_t *P_row_ptr, const uint32_t *P_col_ind, const double *P_val,
                         const double *h_gamma, double *h_lambda, uint32_t stateSize,
                         uint32_t knotPoints, const gbdpcg_config *cfg);

/* Host CSR -> cbtd conversion used by the entry above (dtype-generic via `dtype`). */
int gbdpcg_csr_to_cbtd(int dtype, const uint32_t *row_ptr, const uint32_t *col_ind, const void *val,
                       uint32_t stateSize, uint32_t knotPoints, void *cbtd_out);

/* ---------------------------------------------------------------------------
 * Adjoint (differentiation) — extension.  Given upstream g_lambda = dL/dlambda,
 * solves S g_gamma = g_lambda (S symmetric) with the same PCG and writes
 *   g_gamma = dL/dgamma,  g_S = -1/2 (g_gamma lambda^T + lambda g_gamma^T)
 * restricted to the block-tridiagonal pattern, in cbtd layout (pads zero).
 * forward != 0: first solves S lambda = gamma (lambda in/out as in the plain solve)
 * in the same launch, reusing the on-chip S and Pinv; forward == 0: lambda is input.
 * Device pointers; waits for the stream; returns the forward iteration count
 * (forward != 0) or the adjoint iteration count (forward == 0).
 * ------------------------------------------------------------------------- */
int gbdpcg_adjoint_device_f32(uint32_t stateSize, uint32_t knotPoints, const float *d_S,
                              const float *d_Pinv, const float *d_gamma, float *d_lambda,
                              const float *d_glambda, float *d_ggamma, float *d_gS, int forward,
                              const gbdpcg_config *cfg, gbdpcg_result *result);
int gbdpcg_adjoint_device_f64(uint32_t stateSize, uint32_t knotPoints, const double *d_S,
                              const double *d_Pinv, const double *d_gamma, double *d_lambda,
                              const double *d_glambda, double *d_ggamma, double *d_gS, int forward,
                              const gbdpcg_config *cfg, gbdpcg_result *result);

/* ---------------------------------------------------------------------------
 * Batched solves — extension.  Many independent systems of uneven (n, N) in one
 * persistent launch.  All pointers are device pointers and are bound at plan
 * creation; gbdpcg_batch_run() only launches (no allocation, no memset), so it can
 * be captured in a hipGraph and replayed.
 * ------------------------------------------------------------------------- */
#define GBDPCG_SYS_FORWARD 1 /* solve S lambda = gamma                             */
#define GBDPCG_SYS_ADJOINT 2 /* then (or, without FORWARD: only) the adjoint solve */

typedef struct gbdpcg_system_desc {
  const void *S;       /* cbtd, dtype of the batch                                   */
  const void *Pinv;    /* cbtd preconditioner (required in batches)                  */
  const void *gamma;   /* nN                                                         */
  void *lambda;        /* nN, in/out (warm start in, solution out)                   */
  const void *glambda; /* adjoint only: nN upstream gradient                         */
  void *ggamma;        /* adjoint only: nN output                                    */
  void *gS;            /* adjoint only: 3n^2 N output (cbtd)                         */
  uint32_t n, N;
  double exit_tol;
  uint32_t max_iter;
  uint32_t flags;      /* GBDPCG_SYS_*                                               */
  uint32_t wgs;        /* workgroups for this system, 0 = planner's choice            */
  uint32_t reserved;
} gbdpcg_system_desc;

typedef struct gbdpcg_batch gbdpcg_batch;

/* A plan runs one persistent launch, or with cfg->batch_split one per kernel
 * class present (register rows of n <= 14, rows of n = 28, LDS-resident rows for
 * the others), back to back. */
typedef struct gbdpcg_batch_info {
  uint32_t grid;           /* workgroups of the widest launch                       */
  uint32_t rounds;         /* items per workgroup, summed over the launches         */
  uint32_t wgs_per_cu;     /* co-resident workgroups per CU (smallest launch value)  */
  uint32_t lds_bytes;      /* dynamic LDS per workgroup (largest launch value)      */
  uint32_t systems;
  uint32_t total_items;
  double est_cost;         /* planner's makespan estimate, summed (arbitrary units) */
  uint32_t launches;       /* persistent launches per gbdpcg_batch_run              */
  uint32_t reserved;
} gbdpcg_batch_info;

/* cfg supplies exit_mode, recurrence and stream; tolerance/max_iter are per system.
 * lds_budget_bytes: 0 = the whole LDS of a CU; else the cap per workgroup (the
 * LDS kernel takes more workgroups per system to stay under it; a register
 * layout that needs more is replaced where an alternative exists, else the plan
 * is rejected with GBDPCG_ERR_TOO_LARGE). */
int gbdpcg_batch_create(gbdpcg_batch **out, int dtype, const gbdpcg_system_desc *systems,
                        uint32_t count, const gbdpcg_config *cfg, uint32_t lds_budget_bytes);
int gbdpcg_batch_run(gbdpcg_batch *plan, void *stream);
/* Device array of `count` result records (valid after a run completes). */
const gbdpcg_result *gbdpcg_batch_results_device(const gbdpcg_batch *plan);
/* Copies the result records to host memory (waits for `stream`). */
int gbdpcg_batch_results(gbdpcg_batch *plan, gbdpcg_result *host_out, void *stream);
int gbdpcg_batch_info_get(const gbdpcg_batch *plan, gbdpcg_batch_info *info);
/* Per-system workgroup counts chosen by the planner (count entries). */
int gbdpcg_batch_wgs(const gbdpcg_batch *plan, uint32_t *wgs_out);
int gbdpcg_batch_destroy(gbdpcg_batch *plan);

/* Receding-horizon warm start of every system of a plan, in one launch:
 * lambda[k] <- lambda[k+1] for k < N-1 (lambda[N-1] kept), then, if noise_scale != 0,
 * gamma[i] += noise_scale * z_i with z_i = gbdpcg_counter_randn(seed, system index, i).
 * gamma is modified in place (the descriptor's gamma pointer must be writable).
 * Launch only (no allocation, no memset), so it can be captured with the run. */
int gbdpcg_batch_warm_start(gbdpcg_batch *plan, uint64_t seed, double noise_scale, void *stream);

/* ---------------------------------------------------------------------------
 * Autotuner — extension.  Times candidate layouts (threads per workgroup, knots
 * per workgroup, XCD placement, kernel) on the caller's own data and keeps the
 * fastest.  Choices are cached per process, keyed by (device, dtype, shapes,
 * recurrence, exit mode, flags); if the environment variable GBDPCG_TUNE_CACHE
 * names a file, choices are also read from and appended to it.
 * ------------------------------------------------------------------------- */
typedef struct gbdpcg_tune_report {
  uint32_t candidates;     /* layouts timed (0 when the choice came from the cache)    */
  uint32_t from_cache;     /* 1: cached choice, nothing was timed                      */
  uint32_t knots_per_wg;   /* chosen cap on knots per workgroup                        */
  uint32_t threads;        /* chosen threads per workgroup                             */
  uint32_t wgs;            /* single solve: chosen workgroups; batch: launch grid      */
  uint32_t kernel;         /* GBDPCG_KERNEL_LDS, _REGISTER or _ROWCAST                 */
  double best_us;          /* mean time of one solve / one batch run, chosen layout    */
  double default_us;       /* same for the untuned default layout                      */
} gbdpcg_tune_report;

/* Single solve: times the candidate layouts of (n, N) with cfg's recurrence and exit
 * mode on d_S / d_Pinv / d_gamma (device pointers; lambda0 = 0 in a scratch buffer,
 * the caller's data is not modified) and makes the fastest the default for later
 * solves of this shape on this device whose cfg leaves grid[0], knots_per_wg,
 * threads and kernel automatic.  A choice already cached for the shape is
 * returned without timing (report->from_cache = 1).  With cfg->kernel or
 * cfg->threads set, only the matching candidates are timed and the cache is
 * neither read nor written.  trials = timed runs per candidate (0 = 5). */
int gbdpcg_tune_single(int dtype, uint32_t stateSize, uint32_t knotPoints, const void *d_S, const void *d_Pinv,
                       const void *d_gamma, const gbdpcg_config *cfg, uint32_t trials,
                       gbdpcg_tune_report *report);

/* Batch: builds candidate plans of the same systems (cfg->knots_per_wg,
 * cfg->threads, the register layout and cfg->batch_split varied), runs each on
 * the systems' own data and returns the
 * fastest plan in *out.  The forward systems' lambda is restored to its value at
 * entry; adjoint outputs hold the last trial's values.  Not allowed during graph
 * capture.  trials = timed runs per candidate (0 = 3). */
int gbdpcg_batch_autotune(gbdpcg_batch **out, int dtype, const gbdpcg_system_desc *systems, uint32_t count,
                          const gbdpcg_config *cfg, uint32_t trials, gbdpcg_tune_report *report);

/* ---------------------------------------------------------------------------
 * Preconditioner formation (device) — extension (the documented pcg_solve has no
 * Pinv argument, README.md:42, and empty_pinv = 1 is the default, types.cuh:27).
 * kind = GBDPCG_PRECOND_JACOBI or _SS.  Writes the cbtd Pinv (pads zero).
 * Returns GBDPCG_BREAKDOWN-style error GBDPCG_ERR_INVALID if a diagonal block is
 * singular (checked on device, reported after the stream is synchronised).
 * ------------------------------------------------------------------------- */
int gbdpcg_precond_device(int dtype, uint32_t stateSize, uint32_t knotPoints, const void *d_S,
                          void *d_Pinv, int kind, void *stream);
/* Host version (same definition), for building inputs on the CPU. */
int gbdpcg_precond_host(int dtype, uint32_t stateSize, uint32_t knotPoints, const void *h_S,
                        void *h_Pinv, int kind);

/* ---------------------------------------------------------------------------
 * LQR Schur-complement problems — extension used by tests and benchmarks.
 *   A_k = I + dt*G, G_ij ~ N(0, 0.1^2);  B_k = dt*H, H_ij ~ N(0, 1);
 *   Q_k = diag(U[0.1, 10]);  R_k = diag(U[0.01, 1]);  gamma ~ N(0, 1).
 *   theta_0 = Q_0^-1,  theta_k = A_{k-1} Q_{k-1}^-1 A_{k-1}^T + B_{k-1} R_{k-1}^-1 B_{k-1}^T + Q_k^-1,
 *   S_{k,k-1} = phi_k = -A_{k-1} Q_{k-1}^-1,  S_{k-1,k} = phi_k^T.
 * Draw order per knot k = 0..N-1: G (n*n, row-major), H (n*m, row-major), Q (n), R (m);
 * then gamma (nN).  RNG: xoshiro256** seeded through splitmix64; normals by Box-Muller.
 * Output arrays may be NULL.  A: N*n*n row-major, B: N*n*m row-major, Qd: N*n, Rd: N*m.
 * ------------------------------------------------------------------------- */
int gbdpcg_lqr_generate(int dtype, uint32_t n, uint32_t m, uint32_t N, uint64_t seed, double dt,
                        void *h_S, void *h_gamma, double *A, double *B, double *Qd, double *Rd);
/* Schur complement S from given LQR data (host, double in, dtype out). */
int gbdpcg_lqr_schur_host(int dtype, uint32_t n, uint32_t m, uint32_t N, const double *A,
                          const double *B, const double *Qd, const double *Rd, void *h_S);
/* Same on the device (one workgroup per knot); inputs/outputs are device pointers of
 * the batch dtype (A: N*n*n, B: N*n*m row-major, Qd: N*n, Rd: N*m). */
int gbdpcg_lqr_schur_device(int dtype, uint32_t n, uint32_t m, uint32_t N, const void *d_A,
                            const void *d_B, const void *d_Qd, const void *d_Rd, void *d_S,
                            void *stream);
/* Standard-normal vector from the same generator (count values, scale applied). */
int gbdpcg_randn(int dtype, uint64_t seed, uint64_t count, double scale, void *h_out);
/* Counter-based normals (host), the values gbdpcg_batch_warm_start adds on the device:
 * z_i = sqrt(-2 ln u1) cos(2 pi u2), u1 = (h(2i) >> 11 + 1) 2^-53, u2 = (h(2i+1) >> 11) 2^-53,
 * h(j) = splitmix64(splitmix64(seed) ^ (stream * 0xD1B54A32D192ED03) + j). */
int gbdpcg_counter_randn(int dtype, uint64_t seed, uint64_t stream, uint64_t count, double scale, void *h_out);

/* Warm start for receding-horizon re-solves (device): lambda[k] <- lambda[k+1] for
 * k < N-1 (lambda[N-1] kept), gamma += noise (noise may be NULL). */
int gbdpcg_warm_start_device(int dtype, uint32_t stateSize, uint32_t knotPoints, void *d_lambda,
                             void *d_gamma, const void *d_noise, void *stream);

/* Last error message of the calling thread ("" if none). */
const char *gbdpcg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* GBDPCG_H */
