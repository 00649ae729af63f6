// Internal definitions shared by the device kernels (kernels.hip) and the host
// runtime (api.cpp).  Not part of the public ABI.
#pragma once
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/gbdpcg.h"

namespace gbdpcg {

// Threads per workgroup of the PCG kernel: four wave64s, one per SIMD.
constexpr int kThreads = 256;
// Column chunks per matrix row handled by one lane group (power of two).
constexpr int kColSplit = 4;
// Upper bound of workgroups cooperating on one system (limits the LDS scratch
// that gathers the per-workgroup partial sums).
constexpr int kMaxWgsPerSystem = 1024;
// Polls of one exchange word before a wait is declared failed (about a second).
constexpr uint32_t kSpinLimit = 1u << 21;

// Flags of one system descriptor.
enum : uint32_t {
  KF_FWD = 1u,    // forward solve S lambda = gamma
  KF_ADJ = 2u,    // adjoint solve S g_gamma = g_lambda (+ g_S)
  KF_REL = 4u,    // relative-residual exit test instead of |r'z| < tol
  KF_FUSED = 8u,  // single-reduction (Chronopoulos-Gear) recurrence
  KF_NOXCD = 16u, // never use XCD-local (workgroup-scope) exchange stores
};

// Device-side descriptor of one system (one PCG instance).
struct KSys {
  const void *S;
  const void *Pinv;
  const void *gamma;
  void *lambda;
  const void *glambda;
  void *ggamma;
  void *gS;
  uint64_t *slots;        // exchange words of this system: 2 * G * rec_words
  gbdpcg_result *result;  // result record (may be null)
  int *err;               // launch-wide error word: set to 1 on a timed-out wait (a plain store: it may
                          // be pinned host memory)
  double tol;
  uint32_t n, N, G, max_iter;
  uint32_t flags;
  uint32_t pad_;
};

__host__ __device__ constexpr uint32_t padded_n(uint32_t n) { return (n + 3u) & ~3u; }

// Columns handled by one lane of a kColSplit group (multiple of 4), and the padded
// row stride of the on-chip row-major block rows.
__host__ __device__ constexpr uint32_t col_chunk(uint32_t n) {
  const uint32_t cols = 3u * padded_n(n);
  return ((cols + 4u * kColSplit - 1u) / (4u * kColSplit)) * 4u;
}
__host__ __device__ constexpr uint32_t row_stride(uint32_t n) { return col_chunk(n) * kColSplit; }

// Register kernel: lanes sharing one matrix row, and the columns each holds.
// Two lanes per row halve each lane's matrix registers; one lane per row packs
// twice the knots into a workgroup, which is what n = 28 needs (its batches are
// bound by knots resident per CU): FP32 at 512 threads (~250 VGPRs), FP64 at 256
// threads (one wave per SIMD, matrix chunks partly in AGPRs).  FP64 rows of
// n = 28 at 512 threads take two lanes per row: the paired instance there keeps
// S in LDS (the same 7 knots per workgroup, two waves per SIMD).
__host__ __device__ constexpr uint32_t reg_lpr(uint32_t n, int nt, int tsize) {
  return n <= 2 ? 1u : (n == 28 && ((nt == 512 && tsize == 4) || (nt == 256 && tsize == 8)) ? 1u : 2u);
}
__host__ __device__ constexpr uint32_t reg_chunk(uint32_t n, uint32_t lpr) {
  return 4u * ((3u * padded_n(n) + 4u * lpr - 1u) / (4u * lpr));
}

// Words (64-bit) of one exchange record: kRecScalars scalar slots (3 per wave of
// the publishing workgroup, up to 8 waves) + two edge blocks, each a left edge
// (2 knots of n) + a right edge (2 knots of n); single solves use the first
// block, paired forward + adjoint solves both.  Each value is split into 32-bit
// payload words.
constexpr uint32_t kRecScalars = 24;
__host__ __device__ constexpr uint32_t rec_words(uint32_t n, uint32_t word_per_value) {
  return (kRecScalars + 8u * n) * word_per_value;
}

// Number of T elements of the dynamic LDS used by one workgroup owning K knots.
// Layout (T elements): S rows K*n*RS | Pinv rows (K+2)*n*RS (one ghost knot row
// each side) | 7 ghost-extended vectors of VL = (K+4)*nP + RS (two ghost knots
// each side, plus the read-ahead of the last chunk) | 12 wave partials |
// 3*G gathered partials | 16 control words.
__host__ __device__ inline size_t item_lds_elems(uint32_t n, uint32_t K, uint32_t G) {
  const size_t nP = padded_n(n), RS = row_stride(n);
  const size_t VL = (size_t)(K + 4) * nP + RS;
  return (2 * (size_t)K + 2) * n * RS + 7 * VL + 12 + 3 * (size_t)G + 16;
}

inline size_t item_lds_bytes(uint32_t n, uint32_t K, uint32_t G, size_t tsize) {
  return item_lds_elems(n, K, G) * tsize;
}

// Knot range owned by rank g of G over N knots: [first, last).
__host__ __device__ inline uint32_t knot_begin(uint32_t g, uint32_t G, uint32_t N) {
  return (uint32_t)(((uint64_t)g * N) / G);
}

// Item encoding in the batch schedule: system index (20 bits) | rank (12 bits).
constexpr uint32_t kNoItem = 0xFFFFFFFFu;
__host__ __device__ inline uint32_t make_item(uint32_t sys, uint32_t rank) {
  return (sys << 12) | rank;
}

// Counter-based standard normals shared by the host (gbdpcg_counter_randn) and
// the batched warm-start kernel: splitmix64 (Steele, Lea, Flood 2014) as a hash of
// (key + j), two hashes per value, Box-Muller in double.
__host__ __device__ inline uint64_t smix64(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t counter_key(uint64_t seed, uint64_t stream) {
  return smix64(seed) ^ (stream * 0xd1b54a32d192ed03ULL);
}
__host__ __device__ inline double counter_normal(uint64_t key, uint64_t i) {
  const uint64_t h1 = smix64(key + 2 * i), h2 = smix64(key + 2 * i + 1);
  const double u1 = (double)((h1 >> 11) + 1) * 0x1.0p-53;  // (0, 1]
  const double u2 = (double)(h2 >> 11) * 0x1.0p-53;        // [0, 1)
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

}  // namespace gbdpcg

// ---- host-side helpers implemented next to the kernels ----
#include <hip/hip_runtime_api.h>
namespace gbdpcg {
hipError_t pcg_prepare(int dtype, size_t lds_bytes);
hipError_t pcg_occupancy(int dtype, size_t lds_bytes, int *wgs_per_cu);
hipError_t pcg_launch(int dtype, const KSys &one, const KSys *systems, const uint32_t *items, uint32_t rounds,
                      uint32_t stride, uint32_t grid, size_t lds_bytes, hipStream_t stream);
// aux_kernels.hip
hipError_t fill_u64(uint64_t *p, size_t count, uint64_t value, hipStream_t stream);
hipError_t fill_i32(int *p, size_t count, int value, hipStream_t stream);
hipError_t precond_launch(int dtype, uint32_t n, uint32_t N, const void *S, void *Pinv, int kind, int *err,
                          hipStream_t stream);
hipError_t schur_launch(int dtype, uint32_t n, uint32_t m, uint32_t N, const void *A, const void *B,
                        const void *Qd, const void *Rd, void *S, hipStream_t stream);
hipError_t warm_start_launch(int dtype, uint32_t n, uint32_t N, void *lambda, void *gamma, const void *noise,
                             hipStream_t stream);
hipError_t batch_warm_start_launch(int dtype, const KSys *systems, uint32_t count, uint64_t seed, double scale,
                                   hipStream_t stream);
hipError_t exchange_floor_occupancy(int nt, int *wgs_per_cu);
// out: bit 0 = a wait hit its bound, bit 1 = the rounds used XCD-local stores
hipError_t exchange_floor_launch(int nt, uint64_t *slots, uint32_t G, uint32_t rounds, uint32_t stride,
                                 uint32_t allow_xcd, int *out, hipStream_t stream);
// resource-rate kernels of 256 threads: which 0 = FP32 FMA (out: float[grid]),
// 1 = LDS reads, 2 = HBM reads of `count` 16-byte words at src (out: uint32[grid])
hipError_t rate_launch(int which, uint32_t grid, uint32_t reps, const void *src, size_t count, void *out,
                       hipStream_t stream);
// pcg_reg_kernel.hip (register-resident latency kernel) and its variants.
struct RegVariant {
  bool fused = true;  // fused recurrence; false: reference recurrence (256 threads only)
  bool bc = false;    // row-broadcast layout (pcg_rowcast_kernel.hip, n in {12, 14})
  bool pair = false;  // forward + adjoint items iterated together (pcg_pair_kernel_*.hip)
};
bool reg_supported(int n, bool bc = false);
size_t reg_lds_bytes(int dtype, uint32_t n, uint32_t K, int nt, RegVariant v = {});
uint32_t reg_max_knots(int dtype, uint32_t n, int nt, bool bc = false);
hipError_t reg_prepare(int dtype, int n, int nt, size_t lds, RegVariant v = {});
hipError_t reg_occupancy(int dtype, int n, int nt, size_t lds, int *wgs_per_cu, RegVariant v = {});
hipError_t reg_launch(int dtype, int n, int nt, const KSys &one, const KSys *systems, const uint32_t *items,
                      uint32_t rounds, uint32_t stride, uint32_t grid, size_t lds, hipStream_t stream,
                      RegVariant v = {});
const void *rowcast_kernel_ptr(int dtype, int n, int nt, bool fused);
const void *pair_kernel_ptr_f32(int n, int nt, bool bc);
const void *pair_kernel_ptr_f64(int n, int nt, bool bc);
// pcg_reg_ref_kernel.hip: reference-recurrence kernel for (dtype, n, threads), or null
const void *reg_ref_kernel_ptr(int dtype, int n, int nt);
// tune.cpp: autotuner choice cache (a..d: meaning fixed by the key's kind)
struct TuneChoice {
  uint32_t a, b, c, d;
};
}  // namespace gbdpcg
#include <string>
namespace gbdpcg {
std::string device_tag(int dev);
bool tune_lookup(const std::string &key, TuneChoice *out);
void tune_store(const std::string &key, const TuneChoice &c);
void tune_forget_all();
uint64_t tune_generation();  // changes whenever the cached choices change
}  // namespace gbdpcg
struct gbdpcg_batch;
namespace gbdpcg {
// api.cpp: KF_NOXCD when cfg asks for device-scope exchange stores
// (cfg->exchange_scope == GBDPCG_EXCHANGE_DEVICE) or the environment variable
// GBDPCG_DEVICE_SCOPE_EXCHANGE is set to a non-empty value other than "0" (read
// once); else 0.  cfg may be null (the environment alone decides).
uint32_t exchange_flags(const gbdpcg_config *cfg);
// api.cpp: a forward solve on caller-provided device scratch, no allocation
// (see noalloc.hip).  Scratch layout in 64-bit words: [0, 6) the result record
// when no d_result is given, word kScratchErrWord the error word, records from
// kScratchHeaderWords (256-byte aligned).
constexpr size_t kScratchErrWord = 6;
constexpr size_t kScratchHeaderWords = 32;
int solve_on_scratch(int dtype, uint32_t n, uint32_t N, const void *S, const void *Pinv, const void *gamma,
                     void *lambda, const gbdpcg_config *cfg, gbdpcg_result *d_result, uint64_t *scratch,
                     size_t scratch_words);
// batch.cpp: properties of a built plan used by the tuner
bool batch_uses_registers(const gbdpcg_batch *plan);
int batch_kernel(const gbdpcg_batch *plan);  // GBDPCG_KERNEL_LDS / _REGISTER / _ROWCAST
int batch_threads(const gbdpcg_batch *plan);
}  // namespace gbdpcg
