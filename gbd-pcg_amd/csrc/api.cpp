// C ABI of the solver: drop-in solves, batched plans, adjoint, preconditioner
// and problem utilities.  Host code; launches the kernels of pcg_kernel.hip and
// aux_kernels.hip through the helpers declared in internal.h.
//
// Reference mapping (see include/gbdpcg.h for the full table):
//   host solve     <- solvePCG(h_S, h_Pinv, h_gamma, h_lambda, n, N, cfg)  include/interface.cuh:24-91
//   device solve   <- solvePCG(n, N, d_S, d_Pinv, d_gamma, d_lambda, ...)  include/interface.cuh:94-146
//   occupancy      <- checkPcgOccupancy                                    include/pcg.cuh:23-49
// Differences kept on purpose: errors are returned, never exit(); the host solve
// returns the iteration count (the reference returns 1, interface.cuh:90); n and
// N are runtime values (the reference needs -DSTATE_SIZE/-DKNOT_POINTS,
// include/interface.cuh:112); device buffers are cached, not allocated per call.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "internal.h"

namespace gbdpcg {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

#define HIPCHK(expr)                                                            \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return GBDPCG_ERR_DEVICE;                                                 \
    }                                                                           \
  } while (0)

constexpr uint32_t kMaxIterLimit = 100000;  // keeps a solve's epochs far from wrapping
constexpr int kXcds = 8;                    // gfx950: 8 XCDs, workgroups dealt round-robin

static size_t tsize(int dtype) { return dtype == GBDPCG_F64 ? 8 : 4; }

uint32_t exchange_flags(const gbdpcg_config *cfg) {
  static const uint32_t f = [] {
    const char *e = getenv("GBDPCG_DEVICE_SCOPE_EXCHANGE");
    return (e && *e && strcmp(e, "0") != 0) ? (uint32_t)KF_NOXCD : 0u;
  }();
  return f | ((cfg && cfg->exchange_scope == GBDPCG_EXCHANGE_DEVICE) ? (uint32_t)KF_NOXCD : 0u);
}
static bool valid_dtype(int dtype) { return dtype == GBDPCG_F32 || dtype == GBDPCG_F64; }

// ---------------------------------------------------------------------------
// device properties
// ---------------------------------------------------------------------------
struct DeviceInfo {
  int cus = 0;
  size_t max_lds = 0;
  bool ok = false;
};

// the calling thread's current device (0 when it cannot be read: the call that
// follows then reports the HIP error itself)
static int cur_dev() {
  int d = 0;
  return hipGetDevice(&d) == hipSuccess ? d : 0;
}

static int device_info(int dev, DeviceInfo *out) {
  static std::mutex mu;
  static std::map<int, DeviceInfo> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) {
    *out = it->second;
    return GBDPCG_OK;
  }
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  DeviceInfo di;
  di.cus = prop.multiProcessorCount;
  di.max_lds = prop.sharedMemPerBlock;
  di.ok = true;
  cache[dev] = di;
  *out = di;
  return GBDPCG_OK;
}

static int occupancy(int dtype, size_t lds, int *occ) {
  static std::mutex mu;
  static std::map<std::pair<int, size_t>, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(dtype, lds);
  auto it = cache.find(key);
  if (it != cache.end()) {
    *occ = it->second;
    return GBDPCG_OK;
  }
  HIPCHK(pcg_prepare(dtype, lds));
  int o = 0;
  HIPCHK(pcg_occupancy(dtype, lds, &o));
  cache[key] = o;
  *occ = o;
  return GBDPCG_OK;
}

// ---------------------------------------------------------------------------
// per-(device, stream) workspace of the drop-in calls
// ---------------------------------------------------------------------------
struct SlotBuf {
  uint64_t *p = nullptr;
  size_t words = 0;
  bool dirty = false;  // epochs inconsistent (after a timed-out wait): re-zero before use
};

// Layout of one system's launch
struct Layout {
  uint32_t G = 0, stride = 1, grid = 0;
  size_t lds = 0;
  bool reg = false;    // register-resident kernel
  bool bc = false;     // ... in the row-broadcast layout
  bool pair = false;   // ... running forward and adjoint together
  bool fused = true;   // its recurrence (the LDS kernel takes it from the flags)
  int nt = 256;        // threads per workgroup
};

// A layout chosen for (dtype, n, N, recurrence, pairing, the config fields that
// steer the choice), valid while the tuner's choices are unchanged (gen).
struct LayoutMemo {
  uint64_t k0, k1, k2, gen;
  Layout L;
};

struct Workspace {
  std::mutex mu;  // held by a call from its first enqueue to its last read of the fields below
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<uint64_t, SlotBuf> slots;  // keyed by layout signature
  gbdpcg_result *d_result = nullptr;  // device: record of async solves given no record of their own
  // pinned, device-visible host memory written by the kernels in place (no copy
  // back): the synchronous solves' record and the error words, [0] = an exchange
  // wait timed out, [1] = a singular diagonal block in the preconditioner
  gbdpcg_result *h_result = nullptr;
  int *h_err = nullptr;
  void *d_pinv = nullptr;
  size_t pinv_bytes = 0;
  // pinned, device-visible staging of the host-pointer solves: S | Pinv | gamma |
  // lambda, read and written by the kernels in place
  void *h_stage = nullptr;
  size_t stage_bytes = 0;
  // layouts already chosen for this workspace's calls (see cached_layout)
  std::vector<LayoutMemo> layouts;
};

static std::mutex g_ws_mu;
static std::map<std::pair<int, hipStream_t>, std::unique_ptr<Workspace>> g_ws;

static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

static int get_workspace(int dev, hipStream_t stream, Workspace **out) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto key = std::make_pair(dev, stream);
  auto it = g_ws.find(key);
  if (it != g_ws.end()) {
    *out = it->second.get();
    return GBDPCG_OK;
  }
  if (capturing(stream)) {
    set_error("first solve on a stream must run outside graph capture (workspace allocation)");
    return GBDPCG_ERR_INVALID;
  }
  auto ws = std::make_unique<Workspace>();
  ws->device = dev;
  ws->stream = stream;
  HIPCHK(hipMalloc(&ws->d_result, sizeof(gbdpcg_result)));
  // coherent host memory: never cached on the device, so what the kernels write is
  // in host memory once the stream is synchronised, and host writes are seen by
  // the next launch
  HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ws->h_result), sizeof(gbdpcg_result),
                       hipHostMallocMapped | hipHostMallocCoherent));
  HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ws->h_err), 2 * sizeof(int),
                       hipHostMallocMapped | hipHostMallocCoherent));
  ws->h_err[0] = ws->h_err[1] = 0;
  *out = ws.get();
  g_ws[key] = std::move(ws);
  return GBDPCG_OK;
}

// Exchange-record buffer for one layout; zeroed when created or after a failure.
static int get_slots(Workspace *ws, uint64_t sig, size_t words, uint64_t **out) {
  auto &sb = ws->slots[sig];
  if (sb.p == nullptr) {
    if (capturing(ws->stream)) {
      set_error("first solve of a new size must run outside graph capture (workspace allocation)");
      return GBDPCG_ERR_INVALID;
    }
    HIPCHK(hipMalloc(&sb.p, words * sizeof(uint64_t)));
    sb.words = words;
    HIPCHK(hipMemsetAsync(sb.p, 0, words * sizeof(uint64_t), ws->stream));
  } else if (sb.dirty) {
    HIPCHK(fill_u64(sb.p, sb.words, 0, ws->stream));
    sb.dirty = false;
  }
  *out = sb.p;
  return GBDPCG_OK;
}

static int grow(void **p, size_t *cap, size_t need, hipStream_t s) {
  if (*cap >= need) return GBDPCG_OK;
  if (capturing(s)) {
    set_error("buffer growth inside graph capture; run the same call once outside capture first");
    return GBDPCG_ERR_INVALID;
  }
  if (*p) HIPCHK(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  HIPCHK(hipMalloc(p, need));
  *cap = need;
  return GBDPCG_OK;
}

// Pinned host staging of at least `need` bytes (the caller holds the workspace
// lock, and every earlier use of the buffer has completed: the host-pointer
// solves synchronise before they return).
static int grow_pinned(Workspace *ws, size_t need) {
  if (ws->stage_bytes >= need) return GBDPCG_OK;
  if (ws->h_stage) HIPCHK(hipHostFree(ws->h_stage));
  ws->h_stage = nullptr;
  ws->stage_bytes = 0;
  HIPCHK(hipHostMalloc(&ws->h_stage, need, hipHostMallocMapped | hipHostMallocCoherent));
  ws->stage_bytes = need;
  return GBDPCG_OK;
}

// ---------------------------------------------------------------------------
// workgroup count for a single system
// ---------------------------------------------------------------------------

static uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }


// Latency-oriented choice: about 8 knots per workgroup (the per-iteration cost is
// dominated by the two cross-workgroup exchanges, whose latency grows slowly with
// the number of workgroups, while the on-chip matvec grows with the knots per
// workgroup), all workgroups on one XCD when they fit there.
static int plan_single(int dtype, uint32_t n, uint32_t N, bool fused, uint32_t want_G, uint32_t kcap, Layout *L) {
  DeviceInfo di;
  int rc = device_info(cur_dev(), &di);
  if (rc < 0) return rc;
  const uint32_t maxG = fused ? N / 2 : N;
  if (maxG == 0) {
    set_error("fused recurrence needs at least 2 knots per workgroup");
    return GBDPCG_ERR_INVALID;
  }
  if (!want_G && kcap) want_G = std::min(ceil_div(N, kcap), maxG);
  uint32_t G = want_G ? want_G : std::min<uint32_t>(std::max<uint32_t>(ceil_div(N, 8), 1), 32);
  if (G > maxG) G = maxG;
  while (item_lds_bytes(n, ceil_div(N, G), G, tsize(dtype)) > di.max_lds && G < maxG && !want_G) ++G;
  const size_t lds = item_lds_bytes(n, ceil_div(N, G), G, tsize(dtype));
  if (lds > di.max_lds) {
    set_error("system n=%u N=%u with %u workgroups needs %zu B of LDS per workgroup (max %zu)", n, N, G, lds,
              di.max_lds);
    return GBDPCG_ERR_TOO_LARGE;
  }
  if (G > (uint32_t)kMaxWgsPerSystem) {
    set_error("too many workgroups per system (%u)", G);
    return GBDPCG_ERR_TOO_LARGE;
  }
  int occ = 0;
  rc = occupancy(dtype, lds, &occ);
  if (rc < 0) return rc;
  const uint32_t per_xcd = (uint32_t)(di.cus / kXcds) * (uint32_t)occ;
  const uint32_t total = (uint32_t)di.cus * (uint32_t)occ;
  if (occ <= 0 || G > total) {
    set_error("%u workgroups cannot be co-resident (capacity %u at %zu B LDS)", G, total, lds);
    return GBDPCG_ERR_TOO_LARGE;
  }
  L->G = G;
  L->lds = lds;
  L->stride = G <= per_xcd ? (uint32_t)kXcds : 1u;
  L->grid = G * L->stride;
  return GBDPCG_OK;
}

// Register-resident kernel (either recurrence): reg_lpr(n) lanes per matrix row of
// the owned knots plus one ghost knot each side.  Prefers 256 threads (one wave
// per SIMD: every wave runs the uniform scalar work, so fewer waves cost less)
// with up to 8 knots per workgroup; 512 threads when the knots need it, or when
// only then do all workgroups fit one XCD (shared L2: XCD-local exchanges).
// Returns GBDPCG_ERR_UNSUPPORTED when the shape does not suit this kernel.
static int plan_single_reg(int dtype, uint32_t n, uint32_t N, uint32_t want_G, uint32_t kcap, Layout *L,
                           int only_nt = 0, bool fused = true, bool bc = false, bool pair = false) {
  const RegVariant rv{fused, bc, pair && fused && n <= 14};
  if (!reg_supported((int)n, bc) || N < 4) return GBDPCG_ERR_UNSUPPORTED;
  DeviceInfo di;
  int rc = device_info(cur_dev(), &di);
  if (rc < 0) return rc;
  Layout spread;  // first layout found whose workgroups span several XCDs
  bool have_spread = false;
  for (int nt : {256, 512, 64}) {
    if (only_nt && nt != only_nt) continue;
    if (nt == 64 && (!bc || !only_nt)) continue;  // one-wave ranks: on request (or tuned)
    if (!fused && nt == 512) continue;  // reference-recurrence instances: 64 and 256 threads
    // FP64 rows of n = 28 (and whole FP64 rows in the row-broadcast layout) need
    // more VGPRs than two waves per SIMD leave: one wave per SIMD only
    if (dtype == GBDPCG_F64 && (n > 14 || bc) && nt > 256) continue;
    const uint32_t kmax = reg_max_knots(dtype, n, nt, bc);
    if (kmax < 2) continue;
    uint32_t G = want_G;
    if (!G) {  // about 8 knots (or the cap) per workgroup, never fewer than 2
      const uint32_t target = std::min<uint32_t>(kmax, kcap ? std::max<uint32_t>(kcap, 2u) : 8u);
      G = std::min(ceil_div(N, target), N / 2);
    }
    if (G < 1 || N / G < 2 || ceil_div(N, G) > kmax || G > (uint32_t)kMaxWgsPerSystem) continue;
    const size_t lds = reg_lds_bytes(dtype, n, ceil_div(N, G), nt, rv);
    if (lds > di.max_lds) continue;
    HIPCHK(reg_prepare(dtype, (int)n, nt, lds, rv));
    int occ = 0;
    HIPCHK(reg_occupancy(dtype, (int)n, nt, lds, &occ, rv));
    const uint32_t per_xcd = (uint32_t)(di.cus / kXcds) * (uint32_t)occ;
    if (occ <= 0 || G > (uint32_t)di.cus * (uint32_t)occ) continue;
    Layout c;
    c.G = G;
    c.lds = lds;
    c.reg = true;
    c.bc = bc;
    c.pair = rv.pair;
    c.fused = fused;
    c.nt = nt;
    c.stride = G <= per_xcd ? (uint32_t)kXcds : 1u;
    c.grid = G * c.stride;
    if (c.stride == (uint32_t)kXcds) {
      *L = c;
      return GBDPCG_OK;
    }
    if (!have_spread) {
      spread = c;
      have_spread = true;
    }
  }
  if (!have_spread) return GBDPCG_ERR_UNSUPPORTED;
  *L = spread;
  return GBDPCG_OK;
}

// Layout of a single solve: the autotuner's choice for this shape when the config
// leaves everything automatic, else the planners (register kernel first when the
// recurrence is fused).
static std::string single_key(int dev, int dtype, uint32_t n, uint32_t N, bool fused, int exit_mode) {
  return "single:" + device_tag(dev) + ":" + std::to_string(dtype) + ":" + std::to_string(n) + ":" +
         std::to_string(N) + ":" + std::to_string(fused ? 1 : 0) + ":" + std::to_string(exit_mode);
}

// Rebuilds a layout from its (kernel, threads, workgroups, stride); checks that it
// still fits the device.
static int layout_from(int dtype, uint32_t n, uint32_t N, bool reg, bool bc, int nt, uint32_t G,
                       uint32_t stride, Layout *L, bool fused, bool pair = false) {
  DeviceInfo di;
  int rc = device_info(cur_dev(), &di);
  if (rc < 0) return rc;
  if (G == 0 || G > N || (stride != 1 && stride != (uint32_t)kXcds)) return GBDPCG_ERR_UNSUPPORTED;
  Layout t;
  t.G = G;
  t.stride = stride;
  t.grid = G * stride;
  t.reg = reg;
  t.bc = reg && bc;
  t.pair = reg && fused && pair && n <= 14;
  t.fused = fused;
  const RegVariant rv{fused, t.bc, t.pair};
  t.nt = reg ? nt : kThreads;
  int occ = 0;
  if (reg) {
    if (!reg_supported((int)n, t.bc) || ceil_div(N, G) > reg_max_knots(dtype, n, nt, t.bc) || N / G < 2 ||
        (!fused && nt == 512) || (nt == 64 && (!t.bc || t.pair)) || (t.bc && dtype == GBDPCG_F64 && nt == 512))
      return GBDPCG_ERR_UNSUPPORTED;
    t.lds = reg_lds_bytes(dtype, n, ceil_div(N, G), nt, rv);
    if (t.lds > di.max_lds) return GBDPCG_ERR_UNSUPPORTED;
    HIPCHK(reg_prepare(dtype, (int)n, nt, t.lds, rv));
    HIPCHK(reg_occupancy(dtype, (int)n, nt, t.lds, &occ, rv));
  } else {
    t.lds = item_lds_bytes(n, ceil_div(N, G), G, tsize(dtype));
    if (t.lds > di.max_lds) return GBDPCG_ERR_UNSUPPORTED;
    rc = occupancy(dtype, t.lds, &occ);
    if (rc < 0) return rc;
  }
  const uint32_t per_xcd = (uint32_t)(di.cus / kXcds) * (uint32_t)occ;
  if (occ <= 0 || G > (uint32_t)di.cus * (uint32_t)occ || (stride == (uint32_t)kXcds && G > per_xcd))
    return GBDPCG_ERR_UNSUPPORTED;
  *L = t;
  return GBDPCG_OK;
}

static int choose_layout(int dev, int dtype, uint32_t n, uint32_t N, bool fused, const gbdpcg_config *cfg,
                         Layout *L, bool pair = false) {
  if (cfg->grid[0] == 0 && cfg->knots_per_wg == 0 && cfg->kernel == GBDPCG_KERNEL_AUTO) {
    TuneChoice c;
    if (cfg->threads == 0 && tune_lookup(single_key(dev, dtype, n, N, fused, cfg->exit_mode), &c) &&
        layout_from(dtype, n, N, c.a != 0, c.a == 2, (int)c.b, c.c, c.d, L, fused, pair) == GBDPCG_OK)
      return GBDPCG_OK;
  }
  // automatic: the row-broadcast layout first for FP64 (measured faster there),
  // the column-chunk register layout first for FP32.  cfg->threads asks the
  // register kernels for that many threads per workgroup.
  const int req_nt = (int)cfg->threads;
  int rc = GBDPCG_ERR_UNSUPPORTED;
  if ((cfg->kernel == GBDPCG_KERNEL_AUTO && dtype == GBDPCG_F64) || cfg->kernel == GBDPCG_KERNEL_ROWCAST)
    rc = plan_single_reg(dtype, n, N, cfg->grid[0], cfg->knots_per_wg, L, req_nt, fused, true, pair);
  if (rc == GBDPCG_ERR_UNSUPPORTED && cfg->kernel == GBDPCG_KERNEL_ROWCAST) {
    set_error("row-broadcast kernel needs n in {12,14}, N >= 4 and a fitting layout");
    return GBDPCG_ERR_UNSUPPORTED;
  }
  if (rc == GBDPCG_ERR_UNSUPPORTED && (cfg->kernel == GBDPCG_KERNEL_AUTO || cfg->kernel == GBDPCG_KERNEL_REGISTER))
    rc = plan_single_reg(dtype, n, N, cfg->grid[0], cfg->knots_per_wg, L, req_nt, fused, false, pair);
  if (rc == GBDPCG_ERR_UNSUPPORTED && cfg->kernel == GBDPCG_KERNEL_REGISTER) {
    set_error("register kernel needs n in {2,6,12,14,28}, N >= 4 and a fitting layout");
    return GBDPCG_ERR_UNSUPPORTED;
  }
  if (rc == GBDPCG_ERR_UNSUPPORTED) rc = plan_single(dtype, n, N, fused, cfg->grid[0], cfg->knots_per_wg, L);
  return rc;
}

// Whether the LDS-resident kernel cannot give every workgroup the 2 knots the
// fused recurrence needs: at most N / 2 workgroups, so odd N leaves one of them
// 3 knots, whose S and Phi^-1 rows exceed the LDS of a CU in FP64 from n = 25 on.
static bool lds_fused_unfit(int dtype, uint32_t n, uint32_t N) {
  DeviceInfo di;
  if (N < 2 || device_info(cur_dev(), &di) < 0) return false;
  const uint32_t G = N / 2;
  return item_lds_bytes(n, ceil_div(N, G), G, tsize(dtype)) > di.max_lds;
}

// choose_layout for a workspace's calls, remembered per workspace: a repeated
// solve of the same shape and config skips the tuner-cache key, the device
// queries and the occupancy check (microseconds on every call otherwise).  The
// caller holds ws->mu.
static int cached_layout(Workspace *ws, int dtype, uint32_t n, uint32_t N, bool fused, const gbdpcg_config *cfg,
                         Layout *L, bool pair) {
  const uint64_t k0 = (uint64_t)n | ((uint64_t)N << 32);
  const uint64_t k1 = (uint64_t)cfg->grid[0] | ((uint64_t)cfg->knots_per_wg << 32);
  const uint64_t k2 = (uint64_t)(dtype & 3) | ((uint64_t)fused << 2) | ((uint64_t)pair << 3) |
                      ((uint64_t)(cfg->exit_mode & 3) << 4) | ((uint64_t)(cfg->kernel & 15) << 8) |
                      ((uint64_t)cfg->threads << 16);
  const uint64_t gen = tune_generation();
  for (const LayoutMemo &m : ws->layouts)
    if (m.k0 == k0 && m.k1 == k1 && m.k2 == k2 && m.gen == gen) {
      *L = m.L;
      return GBDPCG_OK;
    }
  const int rc = choose_layout(ws->device, dtype, n, N, fused, cfg, L, pair);
  if (rc < 0) return rc;
  for (LayoutMemo &m : ws->layouts)
    if (m.k0 == k0 && m.k1 == k1 && m.k2 == k2) {
      m.gen = gen;
      m.L = *L;
      return GBDPCG_OK;
    }
  if (ws->layouts.size() >= 32) ws->layouts.erase(ws->layouts.begin());
  ws->layouts.push_back(LayoutMemo{k0, k1, k2, gen, *L});
  return GBDPCG_OK;
}

static int validate_cfg(const gbdpcg_config *cfg) {
  if (!cfg) {
    set_error("null config");
    return GBDPCG_ERR_INVALID;
  }
  if (cfg->max_iter > kMaxIterLimit) {
    set_error("max_iter %u above the limit %u", cfg->max_iter, kMaxIterLimit);
    return GBDPCG_ERR_INVALID;
  }
  if (!(cfg->exit_tol >= 0.0)) {
    set_error("exit_tol must be >= 0");
    return GBDPCG_ERR_INVALID;
  }
  if (cfg->exit_mode != GBDPCG_EXIT_ABS_ETA && cfg->exit_mode != GBDPCG_EXIT_REL_RESID) {
    set_error("bad exit_mode %d", cfg->exit_mode);
    return GBDPCG_ERR_INVALID;
  }
  if (cfg->recurrence != GBDPCG_RECUR_REFERENCE && cfg->recurrence != GBDPCG_RECUR_FUSED) {
    set_error("bad recurrence %d", cfg->recurrence);
    return GBDPCG_ERR_INVALID;
  }
  if (cfg->kernel > GBDPCG_KERNEL_ROWCAST) {
    set_error("bad kernel %u", cfg->kernel);
    return GBDPCG_ERR_INVALID;
  }
  if (cfg->threads != 0 && cfg->threads != 64 && cfg->threads != 256 && cfg->threads != 512) {
    set_error("bad threads %u (0, 64, 256 or 512)", cfg->threads);
    return GBDPCG_ERR_INVALID;
  }
  if (cfg->precond != GBDPCG_PRECOND_JACOBI && cfg->precond != GBDPCG_PRECOND_SS) {
    set_error("bad precond %d", cfg->precond);
    return GBDPCG_ERR_INVALID;
  }
  if (cfg->exchange_scope != GBDPCG_EXCHANGE_AUTO && cfg->exchange_scope != GBDPCG_EXCHANGE_DEVICE) {
    set_error("bad exchange_scope %u", cfg->exchange_scope);
    return GBDPCG_ERR_INVALID;
  }
  return GBDPCG_OK;
}

// The error words live in pinned host memory; read after the stream is
// synchronised, cleared by host stores (no kernel of this workspace is running).
static int check_error_words(Workspace *ws, uint64_t sig) {
  if (ws->h_err[0]) {
    ws->slots[sig].dirty = true;
    ws->h_err[0] = 0;
    set_error("a cross-workgroup exchange wait timed out (workgroups not co-resident?)");
    return GBDPCG_ERR_TIMEOUT;
  }
  if (ws->h_err[1]) {
    ws->h_err[1] = 0;
    set_error("preconditioner: a diagonal block of S is singular");
    return GBDPCG_ERR_INVALID;
  }
  return GBDPCG_OK;
}

static int check_args(int dtype, uint32_t n, uint32_t N, const void *S, const void *gamma, const void *lambda,
                      const void *glambda, const void *ggamma, const void *gS, uint32_t flags,
                      const gbdpcg_config *cfg) {
  int rc = validate_cfg(cfg);
  if (rc < 0) return rc;
  if (!valid_dtype(dtype) || n == 0 || N == 0 || !S || !lambda) {
    set_error("invalid argument (dtype/n/N/S/lambda)");
    return GBDPCG_ERR_INVALID;
  }
  if ((flags & KF_FWD) && !gamma) {
    set_error("gamma is null");
    return GBDPCG_ERR_INVALID;
  }
  if ((flags & KF_ADJ) && (!glambda || !ggamma || !gS)) {
    set_error("adjoint needs glambda, ggamma and gS");
    return GBDPCG_ERR_INVALID;
  }
  return GBDPCG_OK;
}

// Core of every single-system entry point; the caller holds ws->mu from here to
// its last use of the workspace's buffers.  sync: wait for the stream, check the
// error words and return the iteration count (the kernel writes the record
// straight to pinned host memory); else the record goes to d_out (or the
// workspace's device record) and 0 is returned after the launch.
static int solve_locked(Workspace *ws, int dtype, uint32_t n, uint32_t N, const void *S, const void *Pinv,
                        const void *gamma, void *lambda, const void *glambda, void *ggamma, void *gS,
                        uint32_t flags, const gbdpcg_config *cfg, bool sync, gbdpcg_result *h_out,
                        gbdpcg_result *d_out) {
  hipStream_t stream = ws->stream;
  bool fused = cfg->recurrence == GBDPCG_RECUR_FUSED;
  if (fused && N < 2) fused = false;  // one knot: the reference recurrence
  // forward + adjoint on a register kernel (n <= 14): both solves in one loop
  bool pair = fused && (flags & KF_FWD) && (flags & KF_ADJ) && n <= 14;
  Layout L;
  int rc = cached_layout(ws, dtype, n, N, fused, cfg, &L, pair);
  if (rc == GBDPCG_ERR_TOO_LARGE && fused && lds_fused_unfit(dtype, n, N)) {
    // no layout with 2 knots per workgroup: the reference recurrence, which
    // takes one knot per workgroup (batch plans decide the same per system)
    fused = pair = false;
    rc = cached_layout(ws, dtype, n, N, false, cfg, &L, false);
  }
  if (rc < 0) return rc;

  if (Pinv == nullptr) {
    const size_t bytes = (size_t)3 * n * n * N * tsize(dtype);
    rc = grow(&ws->d_pinv, &ws->pinv_bytes, bytes, stream);
    if (rc < 0) return rc;
    HIPCHK(precond_launch(dtype, n, N, S, ws->d_pinv, cfg->precond, ws->h_err + 1, stream));
    Pinv = ws->d_pinv;
  }
  const size_t words = 2 * (size_t)L.G * rec_words(n, (uint32_t)(tsize(dtype) / 4));
  const uint64_t sig = ((uint64_t)dtype) | ((uint64_t)n << 1) | ((uint64_t)L.G << 12) | ((uint64_t)N << 32);
  uint64_t *slots = nullptr;
  rc = get_slots(ws, sig, words, &slots);
  if (rc < 0) return rc;

  KSys one{};
  one.S = S;
  one.Pinv = Pinv;
  one.gamma = gamma;
  one.lambda = lambda;
  one.glambda = glambda;
  one.ggamma = ggamma;
  one.gS = gS;
  one.slots = slots;
  one.result = sync ? ws->h_result : (d_out ? d_out : ws->d_result);
  one.err = ws->h_err;
  one.tol = cfg->exit_tol;
  one.n = n;
  one.N = N;
  one.G = L.G;
  one.max_iter = cfg->max_iter;
  one.flags = flags | (cfg->exit_mode == GBDPCG_EXIT_REL_RESID ? KF_REL : 0u) | (fused ? KF_FUSED : 0u) |
              exchange_flags(cfg);
  if (L.reg) {
    HIPCHK(reg_launch(dtype, (int)n, L.nt, one, nullptr, nullptr, 0, L.stride, L.grid, L.lds, stream,
                      RegVariant{fused, L.bc, L.pair}));
  } else {
    HIPCHK(pcg_launch(dtype, one, nullptr, nullptr, 0, L.stride, L.grid, L.lds, stream));
  }
  if (!sync) return GBDPCG_OK;

  HIPCHK(hipStreamSynchronize(stream));
  rc = check_error_words(ws, sig);
  if (rc < 0) return rc;
  const gbdpcg_result res = *ws->h_result;
  if (h_out) *h_out = res;
  if (res.status == GBDPCG_TIMED_OUT || res.adj_status == GBDPCG_TIMED_OUT) {
    ws->slots[sig].dirty = true;
    set_error("a cross-workgroup exchange wait timed out");
    return GBDPCG_ERR_TIMEOUT;
  }
  return (flags & KF_FWD) ? res.iters : res.adj_iters;
}

// Every single-system device entry point: the workspace of (current device,
// stream) is locked for the whole call, so calls from several host threads on
// one stream run one after the other.
static int solve_device_impl(int dtype, uint32_t n, uint32_t N, const void *S, const void *Pinv, const void *gamma,
                             void *lambda, const void *glambda, void *ggamma, void *gS, uint32_t flags,
                             const gbdpcg_config *cfg, bool sync, gbdpcg_result *h_out, gbdpcg_result *d_out) {
  int rc = check_args(dtype, n, N, S, gamma, lambda, glambda, ggamma, gS, flags, cfg);
  if (rc < 0) return rc;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  Workspace *ws = nullptr;
  rc = get_workspace(dev, static_cast<hipStream_t>(cfg->stream), &ws);
  if (rc < 0) return rc;
  std::lock_guard<std::mutex> lk(ws->mu);
  return solve_locked(ws, dtype, n, N, S, Pinv, gamma, lambda, glambda, ggamma, gS, flags, cfg, sync, h_out, d_out);
}

// The host-pointer drop-in: the inputs are copied into the workspace's pinned
// staging, which the kernels read in place (no device copy of the inputs), and
// the kernel writes lambda there; one stream synchronisation, then lambda is
// copied out.  The workspace stays locked from the first copy in to the last
// copy out.
template <typename T>
static int solve_host_impl(int dtype, const T *h_S, const T *h_Pinv, const T *h_gamma, T *h_lambda, uint32_t n,
                           uint32_t N, const gbdpcg_config *cfg) {
  int rc = validate_cfg(cfg);
  if (rc < 0) return rc;
  if (!h_S || !h_gamma || !h_lambda || n == 0 || N == 0) {
    set_error("invalid argument (null host pointer or zero size)");
    return GBDPCG_ERR_INVALID;
  }
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipStream_t stream = static_cast<hipStream_t>(cfg->stream);
  if (capturing(stream)) {
    set_error("the host-pointer solve synchronises: it cannot run inside graph capture");
    return GBDPCG_ERR_INVALID;
  }
  Workspace *ws = nullptr;
  rc = get_workspace(dev, stream, &ws);
  if (rc < 0) return rc;
  const size_t mat = (size_t)3 * n * n * N, vec = (size_t)n * N;
  std::lock_guard<std::mutex> lk(ws->mu);
  rc = grow_pinned(ws, (2 * mat + 2 * vec) * sizeof(T));
  if (rc < 0) return rc;
  T *s_S = static_cast<T *>(ws->h_stage);
  T *s_P = s_S + mat;
  T *s_g = s_P + mat;
  T *s_l = s_g + vec;
  memcpy(s_S, h_S, mat * sizeof(T));
  if (h_Pinv) memcpy(s_P, h_Pinv, mat * sizeof(T));
  memcpy(s_g, h_gamma, vec * sizeof(T));
  memcpy(s_l, h_lambda, vec * sizeof(T));
  const int iters = solve_locked(ws, dtype, n, N, s_S, h_Pinv ? s_P : nullptr, s_g, s_l, nullptr, nullptr, nullptr,
                                 KF_FWD, cfg, true, nullptr, nullptr);
  if (iters < 0) return iters;
  memcpy(h_lambda, s_l, vec * sizeof(T));
  return iters;
}

template <typename T>
static int solve_csr_impl(int dtype, const uint32_t *Sr, const uint32_t *Sc, const T *Sv, const uint32_t *Pr,
                          const uint32_t *Pc, const T *Pv, const T *h_gamma, T *h_lambda, uint32_t n, uint32_t N,
                          const gbdpcg_config *cfg);

// A forward solve launched on caller-provided device scratch (no workspace, no
// allocation, no memset node): `scratch` holds kScratchHeaderWords words (the
// result record when d_result is null, then the error word) followed by the
// exchange records.  The error word and the records are cleared by a fill
// kernel before the solve, so every call starts from zeroed records.  Used by
// the no-allocation entry point (noalloc.hip).
int solve_on_scratch(int dtype, uint32_t n, uint32_t N, const void *S, const void *Pinv, const void *gamma,
                     void *lambda, const gbdpcg_config *cfg, gbdpcg_result *d_result, uint64_t *scratch,
                     size_t scratch_words) {
  int rc = check_args(dtype, n, N, S, gamma, lambda, nullptr, nullptr, nullptr, KF_FWD, cfg);
  if (rc < 0) return rc;
  if (!Pinv) {
    set_error("the no-allocation solve needs Pinv (forming it would need device memory)");
    return GBDPCG_ERR_INVALID;
  }
  if (!scratch || scratch_words <= kScratchHeaderWords) {
    set_error("no scratch buffer");
    return GBDPCG_ERR_INVALID;
  }
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  bool fused = cfg->recurrence == GBDPCG_RECUR_FUSED;
  if (fused && N < 2) fused = false;
  Layout L;
  rc = choose_layout(dev, dtype, n, N, fused, cfg, &L, false);
  if (rc == GBDPCG_ERR_TOO_LARGE && fused && lds_fused_unfit(dtype, n, N)) {
    fused = false;  // as in solve_locked
    rc = choose_layout(dev, dtype, n, N, false, cfg, &L, false);
  }
  if (rc < 0) return rc;
  const size_t words = 2 * (size_t)L.G * rec_words(n, (uint32_t)(tsize(dtype) / 4));
  if (words > scratch_words - kScratchHeaderWords) {
    set_error("n=%u N=%u needs %zu words of exchange records on %u workgroups; the no-allocation scratch holds "
              "%zu", n, N, words, L.G, scratch_words - kScratchHeaderWords);
    return GBDPCG_ERR_TOO_LARGE;
  }
  hipStream_t stream = static_cast<hipStream_t>(cfg->stream);
  uint64_t *err_word = scratch + kScratchErrWord;
  // the error word and every record word of this layout start at zero
  HIPCHK(fill_u64(err_word, kScratchHeaderWords - kScratchErrWord + words, 0, stream));
  KSys one{};
  one.S = S;
  one.Pinv = Pinv;
  one.gamma = gamma;
  one.lambda = lambda;
  one.slots = scratch + kScratchHeaderWords;
  one.result = d_result ? d_result : reinterpret_cast<gbdpcg_result *>(scratch);
  one.err = reinterpret_cast<int *>(err_word);
  one.tol = cfg->exit_tol;
  one.n = n;
  one.N = N;
  one.G = L.G;
  one.max_iter = cfg->max_iter;
  one.flags = KF_FWD | (cfg->exit_mode == GBDPCG_EXIT_REL_RESID ? KF_REL : 0u) | (fused ? KF_FUSED : 0u) |
              exchange_flags(cfg);
  if (L.reg) {
    HIPCHK(reg_launch(dtype, (int)n, L.nt, one, nullptr, nullptr, 0, L.stride, L.grid, L.lds, stream,
                      RegVariant{fused, L.bc, false}));
  } else {
    HIPCHK(pcg_launch(dtype, one, nullptr, nullptr, 0, L.stride, L.grid, L.lds, stream));
  }
  return GBDPCG_OK;
}

}  // namespace gbdpcg

using namespace gbdpcg;

// ---------------------------------------------------------------------------
// exported entry points
// ---------------------------------------------------------------------------
extern "C" {

const char *gbdpcg_version(void) { return "gbdpcg-mi355x 0.1.0"; }

const char *gbdpcg_last_error(void) { return g_last_error.c_str(); }

void gbdpcg_config_init(gbdpcg_config *cfg) {
  if (!cfg) return;
  memset(cfg, 0, sizeof *cfg);
  cfg->exit_tol = 1e-6;  // include/constants.cuh:16
  cfg->max_iter = 25;    // include/constants.cuh:15
  cfg->grid[0] = 0;      // the reference ignores pcg_grid (interface.cuh:134); 0 = automatic here
  cfg->grid[1] = cfg->grid[2] = 1;
  cfg->block[0] = 64;  // include/constants.cuh:19 (the kernel uses its own 256)
  cfg->block[1] = cfg->block[2] = 1;
  cfg->empty_pinv = 1;  // include/types.cuh:33
  cfg->exit_mode = GBDPCG_EXIT_ABS_ETA;
  cfg->precond = GBDPCG_PRECOND_SS;
  cfg->recurrence = GBDPCG_RECUR_REFERENCE;
  cfg->stream = nullptr;
}

size_t gbdpcg_lds_bytes(int dtype, uint32_t n, uint32_t knots_per_wg, uint32_t wgs) {
  if (!valid_dtype(dtype) || n == 0 || knots_per_wg == 0 || wgs == 0) return 0;
  return item_lds_bytes(n, knots_per_wg, wgs, tsize(dtype));
}

int gbdpcg_check_occupancy(int dtype, uint32_t n, uint32_t N, uint32_t *max_wgs) {
  if (!valid_dtype(dtype) || n == 0 || N == 0) {
    set_error("invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  Layout L;
  int rc = plan_single(dtype, n, N, false, 0, 0, &L);
  if (rc < 0) return rc;
  int occ = 0;
  rc = occupancy(dtype, L.lds, &occ);
  if (rc < 0) return rc;
  DeviceInfo di;
  rc = device_info(cur_dev(), &di);
  if (rc < 0) return rc;
  if (max_wgs) *max_wgs = (uint32_t)(di.cus * occ);
  return GBDPCG_OK;
}

int gbdpcg_exchange_floor(uint32_t wgs, uint32_t threads, uint32_t rounds, double *us_per_round,
                          uint32_t *xcd_local) {
  if (wgs == 0 || wgs > (uint32_t)kMaxWgsPerSystem || (threads != 64 && threads != 256 && threads != 512) ||
      rounds < 20 || rounds > 1000000 || !us_per_round) {
    set_error("gbdpcg_exchange_floor: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  DeviceInfo di;
  int rc = device_info(cur_dev(), &di);
  if (rc < 0) return rc;
  int occ = 0;
  HIPCHK(exchange_floor_occupancy((int)threads, &occ));
  if (occ <= 0 || wgs > (uint32_t)di.cus * (uint32_t)occ) {
    set_error("gbdpcg_exchange_floor: %u workgroups of %u threads are not co-resident", wgs, threads);
    return GBDPCG_ERR_TOO_LARGE;
  }
  // the solvers' placement: all ranks on one XCD when they fit there
  const uint32_t stride = wgs <= (uint32_t)(di.cus / kXcds) * (uint32_t)occ ? (uint32_t)kXcds : 1u;
  const uint32_t allow_xcd = (exchange_flags(nullptr) & KF_NOXCD) ? 0u : 1u;
  uint64_t *slots = nullptr;
  int *out = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  auto cleanup = [&]() {
    if (slots) (void)hipFree(slots);
    if (out) (void)hipFree(out);
    for (hipEvent_t &e : ev)
      if (e) (void)hipEventDestroy(e);
  };
  hipError_t e;
  const uint32_t short_rounds = rounds / 10;
  float t_short = 0.f, t_long = 0.f;
  int h_out = 0;
  if ((e = hipMalloc(&slots, 2 * (size_t)wgs * 4 * sizeof(uint64_t))) != hipSuccess ||
      (e = hipMalloc(&out, sizeof(int))) != hipSuccess ||
      (e = hipMemset(slots, 0, 2 * (size_t)wgs * 4 * sizeof(uint64_t))) != hipSuccess ||
      (e = hipMemset(out, 0, sizeof(int))) != hipSuccess)
    goto fail;
  for (hipEvent_t &x : ev)
    if ((e = hipEventCreate(&x)) != hipSuccess) goto fail;
  // a warm-up launch, then a short and a long run: their difference is the
  // per-round time without the launch and set-up costs
  if ((e = exchange_floor_launch((int)threads, slots, wgs, short_rounds, stride, allow_xcd, out, nullptr)) !=
          hipSuccess ||
      (e = hipEventRecord(ev[0], nullptr)) != hipSuccess ||
      (e = exchange_floor_launch((int)threads, slots, wgs, short_rounds, stride, allow_xcd, out, nullptr)) !=
          hipSuccess ||
      (e = hipEventRecord(ev[1], nullptr)) != hipSuccess || (e = hipEventRecord(ev[2], nullptr)) != hipSuccess ||
      (e = exchange_floor_launch((int)threads, slots, wgs, rounds, stride, allow_xcd, out, nullptr)) != hipSuccess ||
      (e = hipEventRecord(ev[3], nullptr)) != hipSuccess || (e = hipEventSynchronize(ev[3])) != hipSuccess ||
      (e = hipEventElapsedTime(&t_short, ev[0], ev[1])) != hipSuccess ||
      (e = hipEventElapsedTime(&t_long, ev[2], ev[3])) != hipSuccess ||
      (e = hipMemcpy(&h_out, out, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess)
    goto fail;
  cleanup();
  if (h_out & 1) {
    set_error("gbdpcg_exchange_floor: a wait timed out (workgroups not co-resident?)");
    return GBDPCG_ERR_TIMEOUT;
  }
  *us_per_round = (double)(t_long - t_short) * 1000.0 / (double)(rounds - short_rounds);
  if (xcd_local) *xcd_local = (h_out & 2) ? 1u : 0u;
  return GBDPCG_OK;
fail:
  set_error("gbdpcg_exchange_floor: %s", hipGetErrorString(e));
  cleanup();
  return GBDPCG_ERR_DEVICE;
}

int gbdpcg_measure_rates(double *valu_f32_tflops, double *lds_read_gbs, double *hbm_read_gbs) {
  DeviceInfo di;
  int rc = device_info(cur_dev(), &di);
  if (rc < 0) return rc;
  constexpr double kThreads = 256.0;
  const uint32_t cus = (uint32_t)di.cus;
  const uint32_t grid = cus * 16u;      // 16 workgroups of 256 threads per CU: 16 waves per CU
  const size_t hbm_words = (size_t)1 << 27;  // 2 GiB of 16-byte words
  void *buf = nullptr, *out = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  auto cleanup = [&]() {
    if (buf) (void)hipFree(buf);
    if (out) (void)hipFree(out);
    for (hipEvent_t &e : ev)
      if (e) (void)hipEventDestroy(e);
  };
  hipError_t e;
  // time of `reps` rounds with the launch and set-up costs removed: a short and
  // a long run, after a warm-up
  auto timed = [&](int which, uint32_t reps_short, uint32_t reps_long, const void *src, size_t count,
                   double *ms) -> hipError_t {
    hipError_t x;
    float a = 0.f, b = 0.f;
    if ((x = rate_launch(which, grid, reps_short, src, count, out, nullptr)) != hipSuccess ||
        (x = hipEventRecord(ev[0], nullptr)) != hipSuccess ||
        (x = rate_launch(which, grid, reps_short, src, count, out, nullptr)) != hipSuccess ||
        (x = hipEventRecord(ev[1], nullptr)) != hipSuccess ||
        (x = rate_launch(which, grid, reps_long, src, count, out, nullptr)) != hipSuccess ||
        (x = hipEventRecord(ev[2], nullptr)) != hipSuccess || (x = hipEventSynchronize(ev[2])) != hipSuccess ||
        (x = hipEventElapsedTime(&a, ev[0], ev[1])) != hipSuccess ||
        (x = hipEventElapsedTime(&b, ev[1], ev[2])) != hipSuccess)
      return x;
    *ms = (double)b - (double)a;
    return hipSuccess;
  };
  if ((e = hipMalloc(&out, (size_t)grid * sizeof(uint32_t))) != hipSuccess) goto fail;
  for (hipEvent_t &x : ev)
    if ((e = hipEventCreate(&x)) != hipSuccess) goto fail;
  if (valu_f32_tflops) {
    double ms = 0.0;
    const uint32_t rs = 200, rl = 2200;  // 16 x 8 FMAs per lane per rep
    if ((e = timed(0, rs, rl, nullptr, 0, &ms)) != hipSuccess) goto fail;
    *valu_f32_tflops = ms > 0.0 ? (double)grid * kThreads * (rl - rs) * 16.0 * 8.0 * 2.0 / (ms * 1e-3) / 1e12 : 0.0;
  }
  if (lds_read_gbs) {
    double ms = 0.0;
    const uint32_t rs = 500, rl = 5500;  // 4 loads of 16 bytes per lane per rep
    if ((e = timed(1, rs, rl, nullptr, 0, &ms)) != hipSuccess) goto fail;
    *lds_read_gbs = ms > 0.0 ? (double)grid * kThreads * (rl - rs) * 4.0 * 16.0 / (ms * 1e-3) / 1e9 : 0.0;
  }
  if (hbm_read_gbs) {
    double ms = 0.0;
    float t = 0.f;
    if ((e = hipMalloc(&buf, hbm_words * 16)) != hipSuccess || (e = hipMemset(buf, 1, hbm_words * 16)) != hipSuccess)
      goto fail;
    // a warm-up pass, then five timed passes
    if ((e = rate_launch(2, grid, 0, buf, hbm_words, out, nullptr)) != hipSuccess ||
        (e = hipEventRecord(ev[0], nullptr)) != hipSuccess)
      goto fail;
    for (int i = 0; i < 5; ++i)
      if ((e = rate_launch(2, grid, 0, buf, hbm_words, out, nullptr)) != hipSuccess) goto fail;
    if ((e = hipEventRecord(ev[1], nullptr)) != hipSuccess || (e = hipEventSynchronize(ev[1])) != hipSuccess ||
        (e = hipEventElapsedTime(&t, ev[0], ev[1])) != hipSuccess)
      goto fail;
    ms = (double)t / 5.0;
    *hbm_read_gbs = ms > 0.0 ? (double)hbm_words * 16.0 / (ms * 1e-3) / 1e9 : 0.0;
  }
  cleanup();
  return GBDPCG_OK;
fail:
  set_error("gbdpcg_measure_rates: %s", hipGetErrorString(e));
  cleanup();
  return GBDPCG_ERR_DEVICE;
}

int gbdpcg_solve_f32(const float *h_S, const float *h_Pinv, const float *h_gamma, float *h_lambda,
                     uint32_t stateSize, uint32_t knotPoints, const gbdpcg_config *cfg) {
  return solve_host_impl<float>(GBDPCG_F32, h_S, h_Pinv, h_gamma, h_lambda, stateSize, knotPoints, cfg);
}
int gbdpcg_solve_f64(const double *h_S, const double *h_Pinv, const double *h_gamma, double *h_lambda,
                     uint32_t stateSize, uint32_t knotPoints, const gbdpcg_config *cfg) {
  return solve_host_impl<double>(GBDPCG_F64, h_S, h_Pinv, h_gamma, h_lambda, stateSize, knotPoints, cfg);
}

int gbdpcg_solve_device_f32(uint32_t n, uint32_t N, const float *d_S, const float *d_Pinv, const float *d_gamma,
                            float *d_lambda, const gbdpcg_config *cfg, gbdpcg_result *result) {
  return solve_device_impl(GBDPCG_F32, n, N, d_S, d_Pinv, d_gamma, d_lambda, nullptr, nullptr, nullptr, KF_FWD, cfg,
                           true, result, nullptr);
}
int gbdpcg_solve_device_f64(uint32_t n, uint32_t N, const double *d_S, const double *d_Pinv, const double *d_gamma,
                            double *d_lambda, const gbdpcg_config *cfg, gbdpcg_result *result) {
  return solve_device_impl(GBDPCG_F64, n, N, d_S, d_Pinv, d_gamma, d_lambda, nullptr, nullptr, nullptr, KF_FWD, cfg,
                           true, result, nullptr);
}
int gbdpcg_solve_device_async_f32(uint32_t n, uint32_t N, const float *d_S, const float *d_Pinv,
                                  const float *d_gamma, float *d_lambda, const gbdpcg_config *cfg,
                                  gbdpcg_result *d_result) {
  return solve_device_impl(GBDPCG_F32, n, N, d_S, d_Pinv, d_gamma, d_lambda, nullptr, nullptr, nullptr, KF_FWD, cfg,
                           false, nullptr, d_result);
}
int gbdpcg_solve_device_async_f64(uint32_t n, uint32_t N, const double *d_S, const double *d_Pinv,
                                  const double *d_gamma, double *d_lambda, const gbdpcg_config *cfg,
                                  gbdpcg_result *d_result) {
  return solve_device_impl(GBDPCG_F64, n, N, d_S, d_Pinv, d_gamma, d_lambda, nullptr, nullptr, nullptr, KF_FWD, cfg,
                           false, nullptr, d_result);
}

int gbdpcg_adjoint_device_f32(uint32_t n, uint32_t N, const float *d_S, const float *d_Pinv, const float *d_gamma,
                              float *d_lambda, const float *d_glambda, float *d_ggamma, float *d_gS, int forward,
                              const gbdpcg_config *cfg, gbdpcg_result *result) {
  return solve_device_impl(GBDPCG_F32, n, N, d_S, d_Pinv, d_gamma, d_lambda, d_glambda, d_ggamma, d_gS,
                           (forward ? KF_FWD : 0u) | KF_ADJ, cfg, true, result, nullptr);
}
int gbdpcg_adjoint_device_f64(uint32_t n, uint32_t N, const double *d_S, const double *d_Pinv,
                              const double *d_gamma, double *d_lambda, const double *d_glambda, double *d_ggamma,
                              double *d_gS, int forward, const gbdpcg_config *cfg, gbdpcg_result *result) {
  return solve_device_impl(GBDPCG_F64, n, N, d_S, d_Pinv, d_gamma, d_lambda, d_glambda, d_ggamma, d_gS,
                           (forward ? KF_FWD : 0u) | KF_ADJ, cfg, true, result, nullptr);
}

int gbdpcg_solve_csr_f32(const uint32_t *Sr, const uint32_t *Sc, const float *Sv, const uint32_t *Pr,
                         const uint32_t *Pc, const float *Pv, const float *h_gamma, float *h_lambda, uint32_t n,
                         uint32_t N, const gbdpcg_config *cfg) {
  return solve_csr_impl<float>(GBDPCG_F32, Sr, Sc, Sv, Pr, Pc, Pv, h_gamma, h_lambda, n, N, cfg);
}
int gbdpcg_solve_csr_f64(const uint32_t *Sr, const uint32_t *Sc, const double *Sv, const uint32_t *Pr,
                         const uint32_t *Pc, const double *Pv, const double *h_gamma, double *h_lambda, uint32_t n,
                         uint32_t N, const gbdpcg_config *cfg) {
  return solve_csr_impl<double>(GBDPCG_F64, Sr, Sc, Sv, Pr, Pc, Pv, h_gamma, h_lambda, n, N, cfg);
}

int gbdpcg_tune_single(int dtype, uint32_t n, uint32_t N, const void *d_S, const void *d_Pinv, const void *d_gamma,
                       const gbdpcg_config *cfg, uint32_t trials, gbdpcg_tune_report *report) {
  int rc = validate_cfg(cfg);
  if (rc < 0) return rc;
  if (!valid_dtype(dtype) || n == 0 || N == 0 || !d_S || !d_gamma) {
    set_error("gbdpcg_tune_single: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  if (trials == 0) trials = 5;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipStream_t st = static_cast<hipStream_t>(cfg->stream);
  if (capturing(st)) {
    set_error("gbdpcg_tune_single: cannot tune inside graph capture");
    return GBDPCG_ERR_INVALID;
  }
  bool fused = cfg->recurrence == GBDPCG_RECUR_FUSED && N >= 2;
  const size_t ts = tsize(dtype);

  // candidates: the untuned default first, then every (kernel, threads, knots per
  // workgroup, XCD placement) that fits
  std::vector<Layout> cands;
  auto add = [&](const Layout &L) {
    for (const Layout &c : cands)
      if (c.reg == L.reg && c.bc == L.bc && c.nt == L.nt && c.G == L.G && c.stride == L.stride) return;
    cands.push_back(L);
  };
  {
    Layout def;
    rc = GBDPCG_ERR_UNSUPPORTED;
    if ((cfg->kernel == GBDPCG_KERNEL_AUTO && dtype == GBDPCG_F64) || cfg->kernel == GBDPCG_KERNEL_ROWCAST)
      rc = plan_single_reg(dtype, n, N, 0, 0, &def, 0, fused, true);
    if (rc == GBDPCG_ERR_UNSUPPORTED && (cfg->kernel == GBDPCG_KERNEL_AUTO || cfg->kernel == GBDPCG_KERNEL_REGISTER))
      rc = plan_single_reg(dtype, n, N, 0, 0, &def, 0, fused);
    if (rc == GBDPCG_ERR_UNSUPPORTED && (cfg->kernel == GBDPCG_KERNEL_AUTO || cfg->kernel == GBDPCG_KERNEL_LDS))
      rc = plan_single(dtype, n, N, fused, 0, 0, &def);
    if (rc == GBDPCG_ERR_TOO_LARGE && fused && lds_fused_unfit(dtype, n, N)) {
      // solves of this config run the reference iteration (solve_locked): tune
      // those layouts, cached under that recurrence's key, which they look up
      fused = false;
      rc = plan_single(dtype, n, N, false, 0, 0, &def);
    }
    if (rc < 0) return rc;
    add(def);
  }
  // the cache holds the choice for automatic solves: read and written only when
  // the config leaves the kernel and the threads per workgroup open (looked up
  // under the recurrence the default plan settled on)
  const bool open_cfg = cfg->kernel == GBDPCG_KERNEL_AUTO && cfg->threads == 0;
  // a choice already made for this shape (this process, or the cache file) is kept
  if (open_cfg) {
    TuneChoice c;
    Layout L;
    if (tune_lookup(single_key(dev, dtype, n, N, fused, cfg->exit_mode), &c) &&
        layout_from(dtype, n, N, c.a != 0, c.a == 2, (int)c.b, c.c, c.d, &L, fused) == GBDPCG_OK) {
      if (report) {
        gbdpcg_tune_report r{};
        r.from_cache = 1;
        r.knots_per_wg = ceil_div(N, L.G);
        r.threads = (uint32_t)L.nt;
        r.wgs = L.G;
        r.kernel = L.bc ? GBDPCG_KERNEL_ROWCAST : L.reg ? GBDPCG_KERNEL_REGISTER : GBDPCG_KERNEL_LDS;
        *report = r;
      }
      return GBDPCG_OK;
    }
  }

  const uint32_t Ks[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 64};
  // kind 2: row-broadcast register layout, 1: register, 0: LDS
  for (int kind = 2; kind >= 0; --kind) {
    const uint32_t kid = kind == 2 ? GBDPCG_KERNEL_ROWCAST : kind == 1 ? GBDPCG_KERNEL_REGISTER : GBDPCG_KERNEL_LDS;
    if (cfg->kernel != GBDPCG_KERNEL_AUTO && cfg->kernel != kid) continue;
    if (kind == 0 && cfg->threads) continue;  // a thread request names a register kernel
    for (int nt : {256, 512, 64}) {
      if (!kind && nt != kThreads) continue;
      if (nt == 64 && kind != 2) continue;  // one-wave ranks: row-broadcast layout only
      for (uint32_t K : Ks) {
        const uint32_t G = ceil_div(N, K);
        if (G > (uint32_t)kMaxWgsPerSystem || (fused && N / G < 2)) continue;
        for (uint32_t stride : {(uint32_t)kXcds, 1u}) {
          Layout L;
          if (layout_from(dtype, n, N, kind != 0, kind == 2, nt, G, stride, &L, fused) == GBDPCG_OK) add(L);
        }
      }
    }
  }

  // scratch: lambda (zeroed before every run), Pinv if not given, records, result, error words
  size_t max_words = 0;
  for (const Layout &L : cands) max_words = std::max(max_words, 2 * (size_t)L.G * rec_words(n, (uint32_t)(ts / 4)));
  void *lam = nullptr, *pinv = nullptr;
  uint64_t *slots = nullptr;
  gbdpcg_result *res = nullptr;
  int *err = nullptr;
  hipEvent_t ea = nullptr, eb = nullptr;
  auto cleanup = [&]() {
    if (lam) (void)hipFree(lam);
    if (pinv) (void)hipFree(pinv);
    if (slots) (void)hipFree(slots);
    if (res) (void)hipFree(res);
    if (err) (void)hipFree(err);
    if (ea) (void)hipEventDestroy(ea);
    if (eb) (void)hipEventDestroy(eb);
  };
  auto fail = [&](const char *what, hipError_t e) {
    set_error("gbdpcg_tune_single: %s failed: %s", what, hipGetErrorString(e));
    cleanup();
    return GBDPCG_ERR_DEVICE;
  };
  hipError_t e;
  if ((e = hipMalloc(&lam, (size_t)n * N * ts)) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&slots, max_words * sizeof(uint64_t))) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&res, sizeof(gbdpcg_result))) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&err, 2 * sizeof(int))) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipEventCreate(&ea)) != hipSuccess || (e = hipEventCreate(&eb)) != hipSuccess)
    return fail("hipEventCreate", e);
  if ((e = hipMemsetAsync(err, 0, 2 * sizeof(int), st)) != hipSuccess) return fail("hipMemsetAsync", e);
  const void *P = d_Pinv;
  if (!P) {
    if ((e = hipMalloc(&pinv, (size_t)3 * n * n * N * ts)) != hipSuccess) return fail("hipMalloc", e);
    if ((e = precond_launch(dtype, n, N, d_S, pinv, cfg->precond, err + 1, st)) != hipSuccess)
      return fail("preconditioner launch", e);
    P = pinv;
  }

  double best_us = 1e300, default_us = -1.0;
  size_t best = 0;
  uint32_t timed = 0;
  for (size_t ci = 0; ci < cands.size(); ++ci) {
    const Layout &L = cands[ci];
    if ((e = hipMemsetAsync(slots, 0, max_words * sizeof(uint64_t), st)) != hipSuccess)
      return fail("hipMemsetAsync", e);
    KSys one{};
    one.S = d_S;
    one.Pinv = P;
    one.gamma = d_gamma;
    one.lambda = lam;
    one.slots = slots;
    one.result = res;
    one.err = err;
    one.tol = cfg->exit_tol;
    one.n = n;
    one.N = N;
    one.G = L.G;
    one.max_iter = cfg->max_iter;
    one.flags = KF_FWD | (cfg->exit_mode == GBDPCG_EXIT_REL_RESID ? KF_REL : 0u) | (fused ? KF_FUSED : 0u) |
                exchange_flags(cfg);
    double total = 0.0;
    for (uint32_t t = 0; t <= trials; ++t) {
      if ((e = hipMemsetAsync(lam, 0, (size_t)n * N * ts, st)) != hipSuccess) return fail("hipMemsetAsync", e);
      if ((e = hipEventRecord(ea, st)) != hipSuccess) return fail("hipEventRecord", e);
      e = L.reg ? reg_launch(dtype, (int)n, L.nt, one, nullptr, nullptr, 0, L.stride, L.grid, L.lds, st,
                             RegVariant{fused, L.bc, false})
                : pcg_launch(dtype, one, nullptr, nullptr, 0, L.stride, L.grid, L.lds, st);
      if (e != hipSuccess) return fail("solve launch", e);
      if ((e = hipEventRecord(eb, st)) != hipSuccess) return fail("hipEventRecord", e);
      if ((e = hipEventSynchronize(eb)) != hipSuccess) return fail("hipEventSynchronize", e);
      float ms = 0.f;
      if ((e = hipEventElapsedTime(&ms, ea, eb)) != hipSuccess) return fail("hipEventElapsedTime", e);
      if (t > 0) total += ms * 1000.0;
    }
    int herr[2] = {0, 0};
    if ((e = hipMemcpy(herr, err, sizeof herr, hipMemcpyDeviceToHost)) != hipSuccess) return fail("hipMemcpy", e);
    if (herr[1]) {
      cleanup();
      set_error("gbdpcg_tune_single: preconditioner: a diagonal block of S is singular");
      return GBDPCG_ERR_INVALID;
    }
    if (herr[0]) {  // a wait timed out: not a usable layout
      if ((e = hipMemset(err, 0, sizeof(int))) != hipSuccess) return fail("hipMemset", e);
      continue;
    }
    ++timed;
    const double us = total / trials;
    if (ci == 0) default_us = us;
    if (us < best_us) {
      best_us = us;
      best = ci;
    }
  }
  cleanup();
  if (timed == 0) {
    set_error("gbdpcg_tune_single: no candidate layout completed");
    return GBDPCG_ERR_TIMEOUT;
  }
  const Layout &B = cands[best];
  if (open_cfg)
    tune_store(single_key(dev, dtype, n, N, fused, cfg->exit_mode),
               TuneChoice{B.bc ? 2u : B.reg ? 1u : 0u, (uint32_t)B.nt, B.G, B.stride});
  if (report) {
    gbdpcg_tune_report r{};
    r.candidates = timed;
    r.from_cache = 0;
    r.knots_per_wg = ceil_div(N, B.G);
    r.threads = (uint32_t)B.nt;
    r.wgs = B.G;
    r.kernel = B.bc ? GBDPCG_KERNEL_ROWCAST : B.reg ? GBDPCG_KERNEL_REGISTER : GBDPCG_KERNEL_LDS;
    r.best_us = best_us;
    r.default_us = default_us;
    *report = r;
  }
  return GBDPCG_OK;
}

int gbdpcg_precond_device(int dtype, uint32_t n, uint32_t N, const void *d_S, void *d_Pinv, int kind,
                          void *stream) {
  if (!valid_dtype(dtype) || n == 0 || N == 0 || !d_S || !d_Pinv ||
      (kind != GBDPCG_PRECOND_JACOBI && kind != GBDPCG_PRECOND_SS)) {
    set_error("gbdpcg_precond_device: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  DeviceInfo di;
  int rc = device_info(cur_dev(), &di);
  if (rc < 0) return rc;
  if ((7 * (size_t)n * n + n) * tsize(dtype) > di.max_lds) {
    set_error("gbdpcg_precond_device: n=%u too large for on-chip inversion", n);
    return GBDPCG_ERR_TOO_LARGE;
  }
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipStream_t s = static_cast<hipStream_t>(stream);
  Workspace *ws = nullptr;
  rc = get_workspace(dev, s, &ws);
  if (rc < 0) return rc;
  std::lock_guard<std::mutex> lk(ws->mu);
  HIPCHK(precond_launch(dtype, n, N, d_S, d_Pinv, kind, ws->h_err + 1, s));
  HIPCHK(hipStreamSynchronize(s));
  if (ws->h_err[1]) {
    ws->h_err[1] = 0;
    set_error("gbdpcg_precond_device: a diagonal block of S is singular");
    return GBDPCG_ERR_INVALID;
  }
  return GBDPCG_OK;
}

int gbdpcg_lqr_schur_device(int dtype, uint32_t n, uint32_t m, uint32_t N, const void *d_A, const void *d_B,
                            const void *d_Qd, const void *d_Rd, void *d_S, void *stream) {
  if (!valid_dtype(dtype) || n == 0 || N == 0 || !d_A || (m && (!d_B || !d_Rd)) || !d_Qd || !d_S) {
    set_error("gbdpcg_lqr_schur_device: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  HIPCHK(schur_launch(dtype, n, m, N, d_A, d_B, d_Qd, d_Rd, d_S, static_cast<hipStream_t>(stream)));
  return GBDPCG_OK;
}

int gbdpcg_warm_start_device(int dtype, uint32_t n, uint32_t N, void *d_lambda, void *d_gamma, const void *d_noise,
                             void *stream) {
  if (!valid_dtype(dtype) || n == 0 || N == 0 || !d_lambda || (d_noise && !d_gamma)) {
    set_error("gbdpcg_warm_start_device: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  HIPCHK(warm_start_launch(dtype, n, N, d_lambda, d_gamma, d_noise, static_cast<hipStream_t>(stream)));
  return GBDPCG_OK;
}

}  // extern "C"

namespace gbdpcg {
template <typename T>
static int solve_csr_impl(int dtype, const uint32_t *Sr, const uint32_t *Sc, const T *Sv, const uint32_t *Pr,
                          const uint32_t *Pc, const T *Pv, const T *h_gamma, T *h_lambda, uint32_t n, uint32_t N,
                          const gbdpcg_config *cfg) {
  if (!Sr || n == 0 || N == 0) {
    set_error("CSR solve: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  const size_t mat = (size_t)3 * n * n * N;
  std::vector<T> S(mat), P;
  int rc = gbdpcg_csr_to_cbtd(dtype, Sr, Sc, Sv, n, N, S.data());
  if (rc < 0) return rc;
  if (Pr) {
    P.resize(mat);
    rc = gbdpcg_csr_to_cbtd(dtype, Pr, Pc, Pv, n, N, P.data());
    if (rc < 0) return rc;
  }
  return solve_host_impl<T>(dtype, S.data(), Pr ? P.data() : nullptr, h_gamma, h_lambda, n, N, cfg);
}
}  // namespace gbdpcg
