// Batched plans: many independent systems of uneven (n, N) in one persistent
// launch.  The planner picks the workgroups per system, the dynamic LDS size of
// the launch and a gang schedule of (system, rank) items over the co-resident
// workgroups; gbdpcg_batch_run() then only launches.
//
// Scheduling: workgroup b runs on XCD b % 8.  Systems are placed longest-first;
// a system with G ranks takes the G workgroups that become free earliest,
// within one XCD when G fits there (cheaper exchanges), else across XCDs.  Each
// workgroup runs its items in increasing (start time, system) order.  That
// order is the same total order on every workgroup, so the earliest unfinished
// item always has all of its ranks running and the schedule cannot deadlock.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <vector>

#include "internal.h"

namespace gbdpcg {
void set_error(const char *fmt, ...);
}
using namespace gbdpcg;

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return GBDPCG_ERR_DEVICE;                                                           \
    }                                                                                     \
  } while (0)

struct gbdpcg_batch {
  int dtype = 0;
  int device = 0;
  uint32_t count = 0;
  std::vector<gbdpcg_system_desc> sys;
  std::vector<uint32_t> wgs;
  KSys *d_sys = nullptr;
  uint32_t *d_items = nullptr;
  uint64_t *d_slots = nullptr;
  gbdpcg_result *d_results = nullptr;
  int *d_err = nullptr;
  gbdpcg_batch_info info{};
  size_t slot_words = 0;
  bool reg = false;  // register-resident kernel
  bool bc = false;   // ... in the row-broadcast layout
  bool pair = false; // ... with forward + adjoint systems iterated together
  int kn = 0;        // its stateSize specialisation (0 = runtime dispatch)
  int nt = 256;      // threads per workgroup
};

namespace gbdpcg {
bool batch_uses_registers(const gbdpcg_batch *plan) { return plan && plan->reg; }
int batch_kernel(const gbdpcg_batch *plan) {
  if (!plan) return GBDPCG_KERNEL_AUTO;
  return plan->bc ? GBDPCG_KERNEL_ROWCAST : plan->reg ? GBDPCG_KERNEL_REGISTER : GBDPCG_KERNEL_LDS;
}
int batch_threads(const gbdpcg_batch *plan) { return plan ? plan->nt : 0; }
}  // namespace gbdpcg

namespace {

constexpr int kXcds = 8;
constexpr uint32_t kMaxIterLimit = 100000;

size_t tsize(int dtype) { return dtype == GBDPCG_F64 ? 8 : 4; }

// Largest knots-per-workgroup whose LDS fits `budget` (0 if even one does not).
uint32_t max_knots(uint32_t n, uint32_t N, size_t ts, size_t budget) {
  uint32_t best = 0;
  for (uint32_t K = 1; K <= N; ++K) {
    const uint32_t G = (N + K - 1) / K;
    if (item_lds_bytes(n, K, G, ts) <= budget)
      best = K;
    else
      break;
  }
  return best;
}

// Planner's cost of one system (ns, rough): iterations x (exchanges + on-chip work)
// plus the HBM load of its block rows.
// pair: forward + adjoint iterated together (one loop, both matvecs per iteration).
double system_cost(const gbdpcg_system_desc &d, uint32_t G, bool fused, size_t ts, bool pair) {
  const double K = std::ceil((double)d.N / G);
  const double nP = (double)((d.n + 3) & ~3u);
  const double iters = std::min<double>(d.max_iter, 40.0);
  const double passes = ((d.flags & GBDPCG_SYS_FORWARD) ? 1 : 0) + ((d.flags & GBDPCG_SYS_ADJOINT) ? 1 : 0);
  const double exch = (fused ? 1.0 : 2.0) * (G <= 1 ? 250.0 : 900.0 + 12.0 * G);
  const double mv = 2.0 * K * d.n * 3.0 * nP / 32.0 / 2.4 * (ts / 4.0);
  const double load = 2.0 * 3.0 * d.n * d.n * K * ts / 50.0;
  if (pair && passes == 2) return iters * (exch + 2 * mv) + 3 * exch + load + 3000.0;
  return passes * (iters * (exch + mv) + 2 * exch) + load + 3000.0;
}

}  // namespace

extern "C" {

int gbdpcg_batch_create(gbdpcg_batch **out, int dtype, const gbdpcg_system_desc *systems, uint32_t count,
                        const gbdpcg_config *cfg, uint32_t lds_budget_bytes) {
  if (!out || !systems || count == 0 || !cfg || (dtype != GBDPCG_F32 && dtype != GBDPCG_F64)) {
    set_error("gbdpcg_batch_create: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  if (count >= (1u << 20)) {
    set_error("gbdpcg_batch_create: at most 2^20-1 systems per plan");
    return GBDPCG_ERR_TOO_LARGE;
  }
  *out = nullptr;
  const size_t ts = tsize(dtype);
  const bool fused_cfg = cfg->recurrence == GBDPCG_RECUR_FUSED;
  for (uint32_t i = 0; i < count; ++i) {
    const gbdpcg_system_desc &d = systems[i];
    const bool fw = d.flags & GBDPCG_SYS_FORWARD, ad = d.flags & GBDPCG_SYS_ADJOINT;
    if (d.n == 0 || d.N == 0 || !d.S || !d.Pinv || !d.lambda || (!fw && !ad) || (fw && !d.gamma) ||
        (ad && (!d.glambda || !d.ggamma || !d.gS)) || d.max_iter > kMaxIterLimit || !(d.exit_tol >= 0.0)) {
      set_error("gbdpcg_batch_create: system %u is invalid", i);
      return GBDPCG_ERR_INVALID;
    }
  }
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  const size_t max_lds = prop.sharedMemPerBlock;
  const size_t budget = lds_budget_bytes ? std::min<size_t>(lds_budget_bytes, max_lds) : max_lds;

  // kernel: register-resident rows when every system suits it (fused recurrence,
  // stateSize in the specialised set, at least 4 knots), else LDS-resident rows
  bool use_reg = fused_cfg && cfg->kernel != GBDPCG_KERNEL_LDS;
  bool same_n = true;
  bool f64_wide = false;  // FP64 rows of n = 28: only the 256-thread instance holds them
  for (uint32_t i = 0; i < count; ++i) {
    if (!reg_supported((int)systems[i].n) || systems[i].N < 4) use_reg = false;
    if (dtype == GBDPCG_F64 && systems[i].n > 14) f64_wide = true;
    if (systems[i].n != systems[0].n) same_n = false;
  }
  if (f64_wide && !same_n) use_reg = false;
  if (cfg->kernel == GBDPCG_KERNEL_REGISTER && !use_reg) {
    set_error("register kernel needs the fused recurrence, n in {2,6,12,14,28} and N >= 4 for every system");
    return GBDPCG_ERR_UNSUPPORTED;
  }
  // row-broadcast layout (on request; the autotuner tries it): one stateSize in
  // {12, 14} for every system
  const bool bc = use_reg && same_n && reg_supported((int)systems[0].n, true) &&
                  cfg->kernel == GBDPCG_KERNEL_ROWCAST;
  if (cfg->kernel == GBDPCG_KERNEL_ROWCAST && !bc) {
    set_error("row-broadcast kernel needs the fused recurrence, one n in {12,14} and N >= 4 for every system");
    return GBDPCG_ERR_UNSUPPORTED;
  }
  // forward + adjoint systems run both solves in one loop (the paired kernel;
  // other systems of the plan run one pass after the other on it as usual)
  // on rows of n <= 14
  bool any_pair = false, small_n = true;
  for (uint32_t i = 0; i < count; ++i) {
    if ((systems[i].flags & GBDPCG_SYS_FORWARD) && (systems[i].flags & GBDPCG_SYS_ADJOINT)) any_pair = true;
    if (systems[i].n > 14) small_n = false;
  }
  const bool pair = use_reg && any_pair && small_n;
  const RegVariant rv{true, bc, pair};
  // FP64 whole rows (row broadcast) and FP64 rows of n = 28: 256 threads only
  const bool only256 = f64_wide || (bc && dtype == GBDPCG_F64);
  // threads per workgroup: cfg->threads (or the reference-style block[0]) = 256
  // asks for 256 when every system shares n
  const bool want256 = cfg->threads == 256 || (cfg->threads == 0 && cfg->block[0] == 256);
  const int nt = (use_reg && same_n && (want256 || only256)) ? 256 : 512;
  const int kn = same_n ? (int)systems[0].n : 0;

  // workgroups per system and the launch's LDS size
  std::vector<uint32_t> G(count);
  std::vector<uint8_t> fused(count);
  size_t lds = 0;
  // cfg->knots_per_wg caps the knots per workgroup (the autotuner's main knob)
  const uint32_t kcap = cfg->knots_per_wg ? std::max<uint32_t>(cfg->knots_per_wg, 1u) : 0u;
  for (uint32_t i = 0; use_reg && i < count; ++i) {
    const gbdpcg_system_desc &d = systems[i];
    const uint32_t kmax = reg_max_knots(dtype, d.n, nt, bc);
    const uint32_t target = kcap ? std::min(kmax, std::max<uint32_t>(kcap, 2u)) : kmax;
    uint32_t g = d.wgs;
    if (!g) {  // the cap is a target: every workgroup keeps at least 2 knots
      g = (d.N + target - 1) / target;
      g = std::min(g, std::max<uint32_t>(d.N / 2, 1u));
    }
    if (g == 0 || (d.N + g - 1) / g > kmax || d.N / g < 2 || g > (uint32_t)kMaxWgsPerSystem) {
      set_error("system %u: %u workgroups do not suit the register kernel (N=%u, 2..%u knots each)", i, g, d.N,
                kmax);
      return GBDPCG_ERR_TOO_LARGE;
    }
    G[i] = g;
    fused[i] = 1;
    lds = std::max(lds, reg_lds_bytes(dtype, d.n, (d.N + g - 1) / g, nt, rv));
  }
  for (uint32_t i = 0; !use_reg && i < count; ++i) {
    const gbdpcg_system_desc &d = systems[i];
    uint32_t kmax = max_knots(d.n, d.N, ts, budget);
    if (kcap && kmax) kmax = std::min(kmax, kcap);
    if (kmax == 0) {
      set_error("system %u (n=%u) does not fit %zu B of LDS even with one knot per workgroup", i, d.n, budget);
      return GBDPCG_ERR_TOO_LARGE;
    }
    uint32_t g = d.wgs ? d.wgs : (d.N + kmax - 1) / kmax;
    if (g > d.N || (d.N + g - 1) / g > kmax) {
      set_error("system %u: %u workgroups do not fit (N=%u, at most %u knots per workgroup)", i, g, d.N, kmax);
      return GBDPCG_ERR_TOO_LARGE;
    }
    fused[i] = fused_cfg && d.N / g >= 2;
    if (g > (uint32_t)kMaxWgsPerSystem) {
      set_error("system %u needs %u workgroups (max %d)", i, g, kMaxWgsPerSystem);
      return GBDPCG_ERR_TOO_LARGE;
    }
    G[i] = g;
    lds = std::max(lds, item_lds_bytes(d.n, (d.N + g - 1) / g, g, ts));
  }
  lds = (lds + 255) & ~(size_t)255;
  int occ = 0;
  if (use_reg) {
    HIPCHK(reg_prepare(dtype, kn, nt, lds, rv));
    HIPCHK(reg_occupancy(dtype, kn, nt, lds, &occ, rv));
  } else {
    HIPCHK(pcg_prepare(dtype, lds));
    HIPCHK(pcg_occupancy(dtype, lds, &occ));
  }
  if (occ <= 0) {
    set_error("no workgroup fits at %zu B LDS", lds);
    return GBDPCG_ERR_TOO_LARGE;
  }
  const uint32_t grid = (uint32_t)prop.multiProcessorCount * (uint32_t)occ;
  const uint32_t per_xcd = grid / kXcds;
  for (uint32_t i = 0; i < count; ++i) {
    if (G[i] > grid) {
      set_error("system %u needs %u co-resident workgroups (capacity %u)", i, G[i], grid);
      return GBDPCG_ERR_TOO_LARGE;
    }
  }

  // gang list scheduling
  std::vector<double> cost(count);
  for (uint32_t i = 0; i < count; ++i) cost[i] = system_cost(systems[i], G[i], fused[i], ts, pair);
  std::vector<uint32_t> order(count);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  std::vector<double> free_at(grid, 0.0);
  struct Item {
    double start;
    uint32_t sys, rank;
  };
  std::vector<std::vector<Item>> lists(grid);
  std::vector<uint32_t> idx;
  double makespan = 0.0;
  for (uint32_t s : order) {
    const uint32_t g = G[s];
    std::vector<uint32_t> chosen;
    double start = 0.0;
    if (g <= per_xcd) {
      double best = 1e300;
      for (int x = 0; x < kXcds; ++x) {
        idx.clear();
        for (uint32_t j = 0; j < per_xcd; ++j) idx.push_back((uint32_t)x + (uint32_t)kXcds * j);
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return free_at[a] < free_at[b]; });
        const double st = free_at[idx[g - 1]];
        if (st < best) {
          best = st;
          chosen.assign(idx.begin(), idx.begin() + g);
        }
      }
      start = best;
    } else {
      idx.resize(grid);
      std::iota(idx.begin(), idx.end(), 0u);
      std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return free_at[a] < free_at[b]; });
      chosen.assign(idx.begin(), idx.begin() + g);
      start = free_at[idx[g - 1]];
    }
    // ranks in increasing workgroup index (neighbouring ranks stay close)
    std::sort(chosen.begin(), chosen.end());
    for (uint32_t r = 0; r < g; ++r) {
      lists[chosen[r]].push_back({start, s, r});
      free_at[chosen[r]] = start + cost[s];
    }
    makespan = std::max(makespan, start + cost[s]);
  }
  uint32_t rounds = 0, total_items = 0;
  for (auto &l : lists) {
    std::stable_sort(l.begin(), l.end(), [](const Item &a, const Item &b) {
      return a.start < b.start || (a.start == b.start && a.sys < b.sys);
    });
    rounds = std::max<uint32_t>(rounds, (uint32_t)l.size());
    total_items += (uint32_t)l.size();
  }
  std::vector<uint32_t> items((size_t)rounds * grid, kNoItem);
  for (uint32_t b = 0; b < grid; ++b)
    for (size_t r = 0; r < lists[b].size(); ++r) items[r * grid + b] = make_item(lists[b][r].sys, lists[b][r].rank);

  // device state
  auto plan = new gbdpcg_batch();
  plan->dtype = dtype;
  plan->device = dev;
  plan->count = count;
  plan->sys.assign(systems, systems + count);
  plan->wgs = G;
  std::vector<size_t> slot_off(count);
  size_t words = 0;
  const uint32_t wpv = (uint32_t)(ts / 4);
  for (uint32_t i = 0; i < count; ++i) {
    slot_off[i] = words;
    words += 2 * (size_t)G[i] * rec_words(systems[i].n, wpv);
  }
  plan->slot_words = words;
  auto fail = [&](int rc) {
    gbdpcg_batch_destroy(plan);
    return rc;
  };
  hipError_t e;
  if ((e = hipMalloc(&plan->d_slots, words * sizeof(uint64_t))) != hipSuccess ||
      (e = hipMalloc(&plan->d_sys, count * sizeof(KSys))) != hipSuccess ||
      (e = hipMalloc(&plan->d_items, std::max<size_t>(items.size(), 1) * sizeof(uint32_t))) != hipSuccess ||
      (e = hipMalloc(&plan->d_results, count * sizeof(gbdpcg_result))) != hipSuccess ||
      (e = hipMalloc(&plan->d_err, sizeof(int))) != hipSuccess) {
    set_error("gbdpcg_batch_create: hipMalloc failed: %s", hipGetErrorString(e));
    return fail(GBDPCG_ERR_NOMEM);
  }
  std::vector<KSys> ks(count);
  for (uint32_t i = 0; i < count; ++i) {
    const gbdpcg_system_desc &d = systems[i];
    KSys &k = ks[i];
    memset(&k, 0, sizeof k);
    k.S = d.S;
    k.Pinv = d.Pinv;
    k.gamma = d.gamma;
    k.lambda = d.lambda;
    k.glambda = d.glambda;
    k.ggamma = d.ggamma;
    k.gS = d.gS;
    k.slots = plan->d_slots + slot_off[i];
    k.result = plan->d_results + i;
    k.err = plan->d_err;
    k.tol = d.exit_tol;
    k.n = d.n;
    k.N = d.N;
    k.G = G[i];
    k.max_iter = d.max_iter;
    k.flags = ((d.flags & GBDPCG_SYS_FORWARD) ? KF_FWD : 0u) | ((d.flags & GBDPCG_SYS_ADJOINT) ? KF_ADJ : 0u) |
              (cfg->exit_mode == GBDPCG_EXIT_REL_RESID ? KF_REL : 0u) | (fused[i] ? KF_FUSED : 0u) |
              exchange_flags();
  }
  if ((e = hipMemset(plan->d_slots, 0, words * sizeof(uint64_t))) != hipSuccess ||
      (e = hipMemset(plan->d_results, 0, count * sizeof(gbdpcg_result))) != hipSuccess ||
      (e = hipMemset(plan->d_err, 0, sizeof(int))) != hipSuccess ||
      (e = hipMemcpy(plan->d_sys, ks.data(), count * sizeof(KSys), hipMemcpyHostToDevice)) != hipSuccess ||
      (!items.empty() &&
       (e = hipMemcpy(plan->d_items, items.data(), items.size() * sizeof(uint32_t), hipMemcpyHostToDevice)) !=
           hipSuccess)) {
    set_error("gbdpcg_batch_create: upload failed: %s", hipGetErrorString(e));
    return fail(GBDPCG_ERR_DEVICE);
  }
  plan->info.grid = grid;
  plan->info.rounds = rounds;
  plan->info.wgs_per_cu = (uint32_t)occ;
  plan->info.lds_bytes = (uint32_t)lds;
  plan->reg = use_reg;
  plan->bc = bc;
  plan->pair = pair;
  plan->kn = kn;
  plan->nt = use_reg ? nt : kThreads;
  plan->info.systems = count;
  plan->info.total_items = total_items;
  plan->info.est_cost = makespan;
  *out = plan;
  return GBDPCG_OK;
}

int gbdpcg_batch_run(gbdpcg_batch *plan, void *stream) {
  if (!plan) {
    set_error("gbdpcg_batch_run: null plan");
    return GBDPCG_ERR_INVALID;
  }
  KSys dummy;
  memset(&dummy, 0, sizeof dummy);
  if (plan->reg) {
    HIPCHK(reg_launch(plan->dtype, plan->kn, plan->nt, dummy, plan->d_sys, plan->d_items, plan->info.rounds, 1,
                      plan->info.grid, plan->info.lds_bytes, static_cast<hipStream_t>(stream),
                      RegVariant{true, plan->bc, plan->pair}));
    return GBDPCG_OK;
  }
  HIPCHK(pcg_launch(plan->dtype, dummy, plan->d_sys, plan->d_items, plan->info.rounds, 1, plan->info.grid,
                    plan->info.lds_bytes, static_cast<hipStream_t>(stream)));
  return GBDPCG_OK;
}

const gbdpcg_result *gbdpcg_batch_results_device(const gbdpcg_batch *plan) {
  return plan ? plan->d_results : nullptr;
}

int gbdpcg_batch_results(gbdpcg_batch *plan, gbdpcg_result *host_out, void *stream) {
  if (!plan || !host_out) {
    set_error("gbdpcg_batch_results: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  int err = 0;
  HIPCHK(hipMemcpyAsync(host_out, plan->d_results, plan->count * sizeof(gbdpcg_result), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&err, plan->d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  bool timed_out = err != 0;
  for (uint32_t i = 0; i < plan->count; ++i)
    timed_out = timed_out || host_out[i].status == GBDPCG_TIMED_OUT || host_out[i].adj_status == GBDPCG_TIMED_OUT;
  if (timed_out) {
    // epochs are no longer consistent: restart them from zero
    HIPCHK(hipMemsetAsync(plan->d_slots, 0, plan->slot_words * sizeof(uint64_t), s));
    HIPCHK(hipMemsetAsync(plan->d_err, 0, sizeof(int), s));
    HIPCHK(hipStreamSynchronize(s));
    set_error("a cross-workgroup exchange wait timed out");
    return GBDPCG_ERR_TIMEOUT;
  }
  return GBDPCG_OK;
}

int gbdpcg_batch_info_get(const gbdpcg_batch *plan, gbdpcg_batch_info *info) {
  if (!plan || !info) {
    set_error("gbdpcg_batch_info_get: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  *info = plan->info;
  return GBDPCG_OK;
}

int gbdpcg_batch_wgs(const gbdpcg_batch *plan, uint32_t *wgs_out) {
  if (!plan || !wgs_out) {
    set_error("gbdpcg_batch_wgs: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  std::copy(plan->wgs.begin(), plan->wgs.end(), wgs_out);
  return GBDPCG_OK;
}

int gbdpcg_batch_warm_start(gbdpcg_batch *plan, uint64_t seed, double noise_scale, void *stream) {
  if (!plan || !(noise_scale == noise_scale)) {
    set_error("gbdpcg_batch_warm_start: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  HIPCHK(batch_warm_start_launch(plan->dtype, plan->d_sys, plan->count, seed, noise_scale,
                                 static_cast<hipStream_t>(stream)));
  return GBDPCG_OK;
}

int gbdpcg_batch_destroy(gbdpcg_batch *plan) {
  if (!plan) return GBDPCG_OK;
  if (plan->d_slots) hipFree(plan->d_slots);
  if (plan->d_sys) hipFree(plan->d_sys);
  if (plan->d_items) hipFree(plan->d_items);
  if (plan->d_results) hipFree(plan->d_results);
  if (plan->d_err) hipFree(plan->d_err);
  delete plan;
  return GBDPCG_OK;
}

}  // extern "C"
