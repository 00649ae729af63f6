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

// One persistent launch of a plan: a kernel class and its schedule.  A plan runs
// its launches back to back on one stream (still one capturable call).
struct BatchLaunch {
  bool reg = false;    // register-resident kernel (else LDS-resident)
  bool bc = false;     // ... in the row-broadcast layout
  bool pair = false;   // ... with forward + adjoint systems iterated together
  int kn = 0;          // stateSize specialisation (0 = runtime dispatch)
  int nt = 256;        // threads per workgroup
  uint32_t grid = 0, rounds = 0, occ = 0, items = 0, systems = 0;
  size_t lds = 0;
  size_t item_off = 0;  // offset of this launch's rounds x grid items in d_items
  double makespan = 0.0;
};

struct gbdpcg_batch {
  int dtype = 0;
  int device = 0;
  uint32_t count = 0;
  std::vector<gbdpcg_system_desc> sys;
  std::vector<uint32_t> wgs;
  std::vector<BatchLaunch> launches;
  KSys *d_sys = nullptr;
  uint32_t *d_items = nullptr;
  uint64_t *d_slots = nullptr;
  gbdpcg_result *d_results = nullptr;
  int *d_err = nullptr;
  gbdpcg_batch_info info{};
  size_t slot_words = 0;
};

namespace {
// the launch holding the most systems (what the tuner reports)
const BatchLaunch *primary(const gbdpcg_batch *plan) {
  const BatchLaunch *b = nullptr;
  for (const BatchLaunch &L : plan->launches)
    if (!b || L.systems > b->systems) b = &L;
  return b;
}
}  // namespace

namespace gbdpcg {
bool batch_uses_registers(const gbdpcg_batch *plan) {
  const BatchLaunch *b = plan ? primary(plan) : nullptr;
  return b && b->reg;
}
int batch_kernel(const gbdpcg_batch *plan) {
  const BatchLaunch *b = plan ? primary(plan) : nullptr;
  if (!b) return GBDPCG_KERNEL_AUTO;
  return b->bc ? GBDPCG_KERNEL_ROWCAST : b->reg ? GBDPCG_KERNEL_REGISTER : GBDPCG_KERNEL_LDS;
}
int batch_threads(const gbdpcg_batch *plan) {
  const BatchLaunch *b = plan ? primary(plan) : nullptr;
  return b ? b->nt : 0;
}
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

// Paired instances for rows of n = 28 (one lane per row): FP32 at 512 threads,
// FP64 at 256.
// Paired instances for rows of n = 28: FP32 at 512 threads; FP64 at 256 (one
// lane per row) and, for launches whose systems are all forward + adjoint, at
// 512 (S in LDS; that instance runs paired items only).
bool pair28(int dtype, int nt, bool all_pair) {
  return dtype == GBDPCG_F64 ? (nt == 256 || (nt == 512 && all_pair)) : nt == 512;
}

// Whether the FP64 n = 28 paired instance with S in LDS (512 threads, up to 7
// knots per workgroup) fits an LDS budget per workgroup.
bool s_lds_fits(size_t budget) {
  const size_t need = reg_lds_bytes(GBDPCG_F64, 28, 7, 512, RegVariant{true, false, true});
  return ((need + 255) & ~(size_t)255) <= budget;
}

bool all_paired(const gbdpcg_system_desc *systems, const std::vector<uint32_t> &idx) {
  for (uint32_t i : idx)
    if (!((systems[i].flags & GBDPCG_SYS_FORWARD) && (systems[i].flags & GBDPCG_SYS_ADJOINT))) return false;
  return true;
}

// Plans one launch over the systems `idx` (all of one kernel class): workgroups
// per system (G, fused), the launch's LDS size, co-resident grid and a gang list
// schedule, whose items are appended to `items`.
int plan_launch(int dtype, const gbdpcg_system_desc *systems, const std::vector<uint32_t> &idx,
                const gbdpcg_config *cfg, const hipDeviceProp_t &prop, size_t budget, BatchLaunch &L,
                std::vector<uint32_t> &G, std::vector<uint8_t> &fused, std::vector<uint32_t> &items) {
  const size_t ts = tsize(dtype);
  const bool fused_cfg = cfg->recurrence == GBDPCG_RECUR_FUSED;
  const RegVariant rv{true, L.bc, L.pair};
  size_t lds = 0;
  // cfg->knots_per_wg caps the knots per workgroup (the autotuner's main knob)
  const uint32_t kcap = cfg->knots_per_wg ? std::max<uint32_t>(cfg->knots_per_wg, 1u) : 0u;
  for (uint32_t i : idx) {
    const gbdpcg_system_desc &d = systems[i];
    if (L.reg) {
      const uint32_t kmax = reg_max_knots(dtype, d.n, L.nt, L.bc);
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
      lds = std::max(lds, reg_lds_bytes(dtype, d.n, (d.N + g - 1) / g, L.nt, rv));
    } else {
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
      if (g > (uint32_t)kMaxWgsPerSystem) {
        set_error("system %u needs %u workgroups (max %d)", i, g, kMaxWgsPerSystem);
        return GBDPCG_ERR_TOO_LARGE;
      }
      fused[i] = fused_cfg && d.N / g >= 2;
      G[i] = g;
      lds = std::max(lds, item_lds_bytes(d.n, (d.N + g - 1) / g, g, ts));
    }
  }
  lds = (lds + 255) & ~(size_t)255;
  if (L.reg && lds > budget) {
    set_error("the register kernel's layout needs %zu B of LDS per workgroup, above the budget of %zu B", lds,
              budget);
    return GBDPCG_ERR_TOO_LARGE;
  }
  int occ = 0;
  if (L.reg) {
    HIPCHK(reg_prepare(dtype, L.kn, L.nt, lds, rv));
    HIPCHK(reg_occupancy(dtype, L.kn, L.nt, lds, &occ, rv));
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
  for (uint32_t i : idx) {
    if (G[i] > grid) {
      set_error("system %u needs %u co-resident workgroups (capacity %u)", i, G[i], grid);
      return GBDPCG_ERR_TOO_LARGE;
    }
  }

  // gang list scheduling
  std::vector<double> cost(idx.size());
  for (size_t q = 0; q < idx.size(); ++q) cost[q] = system_cost(systems[idx[q]], G[idx[q]], fused[idx[q]], ts, L.pair);
  std::vector<uint32_t> order(idx.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  std::vector<double> free_at(grid, 0.0);
  struct Item {
    double start;
    uint32_t sys, rank;
  };
  std::vector<std::vector<Item>> lists(grid);
  std::vector<uint32_t> cand;
  double makespan = 0.0;
  for (uint32_t q : order) {
    const uint32_t s = idx[q];
    const uint32_t g = G[s];
    std::vector<uint32_t> chosen;
    double start = 0.0;
    if (g <= per_xcd) {
      double best = 1e300;
      for (int x = 0; x < kXcds; ++x) {
        cand.clear();
        for (uint32_t j = 0; j < per_xcd; ++j) cand.push_back((uint32_t)x + (uint32_t)kXcds * j);
        std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return free_at[a] < free_at[b]; });
        const double st = free_at[cand[g - 1]];
        if (st < best) {
          best = st;
          chosen.assign(cand.begin(), cand.begin() + g);
        }
      }
      start = best;
    } else {
      cand.resize(grid);
      std::iota(cand.begin(), cand.end(), 0u);
      std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return free_at[a] < free_at[b]; });
      chosen.assign(cand.begin(), cand.begin() + g);
      start = free_at[cand[g - 1]];
    }
    // ranks in increasing workgroup index (neighbouring ranks stay close)
    std::sort(chosen.begin(), chosen.end());
    for (uint32_t r = 0; r < g; ++r) {
      lists[chosen[r]].push_back({start, s, r});
      free_at[chosen[r]] = start + cost[q];
    }
    makespan = std::max(makespan, start + cost[q]);
  }
  uint32_t rounds = 0, total_items = 0;
  for (auto &l : lists) {
    std::stable_sort(l.begin(), l.end(), [](const Item &a, const Item &b) {
      return a.start < b.start || (a.start == b.start && a.sys < b.sys);
    });
    rounds = std::max<uint32_t>(rounds, (uint32_t)l.size());
    total_items += (uint32_t)l.size();
  }
  L.item_off = items.size();
  items.resize(items.size() + (size_t)rounds * grid, kNoItem);
  for (uint32_t b = 0; b < grid; ++b)
    for (size_t r = 0; r < lists[b].size(); ++r)
      items[L.item_off + r * grid + b] = make_item(lists[b][r].sys, lists[b][r].rank);
  L.grid = grid;
  L.rounds = rounds;
  L.occ = (uint32_t)occ;
  L.lds = lds;
  L.items = total_items;
  L.systems = (uint32_t)idx.size();
  L.makespan = makespan;
  return GBDPCG_OK;
}

}  // namespace

extern "C" {

int gbdpcg_batch_create(gbdpcg_batch **out, int dtype, const gbdpcg_system_desc *systems, uint32_t count,
                        const gbdpcg_config *cfg, uint32_t lds_budget_bytes) {
  if (!out || !systems || count == 0 || !cfg || (dtype != GBDPCG_F32 && dtype != GBDPCG_F64) ||
      (cfg->exchange_scope != GBDPCG_EXCHANGE_AUTO && cfg->exchange_scope != GBDPCG_EXCHANGE_DEVICE)) {
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

  // Kernel choice.  Register-resident rows take the fused recurrence, the
  // specialised stateSizes and N >= 4; the LDS-resident kernel takes the rest.
  // By default a plan is one persistent launch: the register kernel when every
  // system suits it (the runtime-n instance when stateSizes differ), else the
  // LDS kernel for all.  cfg->batch_split asks for one launch per kernel class
  // present instead: rows of n <= 14 (their runtime-n instance then needs the
  // registers of those sizes only: two 512-thread workgroups per CU), rows of
  // n = 28, and the LDS kernel for systems the register kernels do not take.
  const bool reg_cfg = fused_cfg && cfg->kernel != GBDPCG_KERNEL_LDS;
  const bool split = cfg->batch_split != 0;
  std::vector<uint32_t> small, wide, lds_sys;
  for (uint32_t i = 0; i < count; ++i) {
    const gbdpcg_system_desc &d = systems[i];
    if (reg_cfg && reg_supported((int)d.n) && d.N >= 4)
      (d.n > 14 ? wide : small).push_back(i);
    else
      lds_sys.push_back(i);
  }
  auto same_n = [&](const std::vector<uint32_t> &v) {
    for (uint32_t i : v)
      if (systems[i].n != systems[v[0]].n) return false;
    return true;
  };
  if (!split) {
    // one launch: FP64 rows of n = 28 need the 256-thread instance, which has no
    // runtime-n form, so mixed FP64 plans holding them go to the LDS kernel
    std::vector<uint32_t> all(count);
    std::iota(all.begin(), all.end(), 0u);
    const bool f64_wide_mixed = dtype == GBDPCG_F64 && !wide.empty() && !same_n(all);
    if (lds_sys.empty() && !f64_wide_mixed) {
      small = all;
      wide.clear();
    } else {
      lds_sys = all;
      small.clear();
      wide.clear();
    }
  }
  if (cfg->kernel == GBDPCG_KERNEL_REGISTER && !lds_sys.empty()) {
    set_error("register kernel needs the fused recurrence, n in {2,6,12,14,28} and N >= 4 for every system%s",
              split ? "" : " (FP64 plans mixing n = 28 with other sizes: set batch_split)");
    return GBDPCG_ERR_UNSUPPORTED;
  }
  // row-broadcast layout (on request; the autotuner tries it): one stateSize in
  // {12, 14} for every system
  const bool bc = cfg->kernel == GBDPCG_KERNEL_ROWCAST;
  if (bc && (!lds_sys.empty() || !wide.empty() || !same_n(small) || !reg_supported((int)systems[0].n, true))) {
    set_error("row-broadcast kernel needs the fused recurrence, one n in {12,14} and N >= 4 for every system");
    return GBDPCG_ERR_UNSUPPORTED;
  }
  // threads per workgroup: cfg->threads (or the reference-style block[0]) = 256
  // asks for 256 when a launch's systems share n
  const bool want256 = cfg->threads == 256 || (cfg->threads == 0 && cfg->block[0] == 256);

  std::vector<BatchLaunch> launches;
  std::vector<std::vector<uint32_t>> members;
  if (!small.empty()) {
    BatchLaunch L;
    L.reg = true;
    L.bc = bc;
    const bool one_n = same_n(small);
    bool has_wide = false;
    for (uint32_t i : small)
      if (systems[i].n > 14) has_wide = true;
    // runtime-n instances: 0 = every specialised size, -1 = sizes up to 14
    L.kn = one_n ? (int)systems[small[0]].n : has_wide ? 0 : -1;
    // FP64 whole rows (row broadcast) and FP64 rows of n = 28: 256 threads, but
    // for a launch of forward + adjoint systems of n = 28 only, 512 (S in LDS)
    const bool pair_only = all_paired(systems, small);
    const bool f64_pair28 = dtype == GBDPCG_F64 && one_n && systems[small[0]].n == 28 && pair_only && !bc &&
                            s_lds_fits(budget);
    const bool only256 = dtype == GBDPCG_F64 && (bc || has_wide) && !f64_pair28;
    L.nt = (one_n && (want256 || only256)) ? 256 : 512;
    // forward + adjoint systems run both solves in one loop (the paired kernel;
    // other systems of the launch run one pass after the other on it): rows of
    // n <= 14, and rows of n = 28 on their one-lane-per-row instances
    for (uint32_t i : small)
      if ((systems[i].flags & GBDPCG_SYS_FORWARD) && (systems[i].flags & GBDPCG_SYS_ADJOINT)) L.pair = true;
    if (has_wide && !(one_n && pair28(dtype, L.nt, pair_only))) L.pair = false;
    launches.push_back(L);
    members.push_back(small);
  }
  if (!wide.empty()) {
    BatchLaunch L;
    L.reg = true;
    L.kn = 28;
    const bool pair_only = all_paired(systems, wide) && (dtype != GBDPCG_F64 || s_lds_fits(budget));
    L.nt = (want256 || (dtype == GBDPCG_F64 && !pair_only)) ? 256 : 512;
    for (uint32_t i : wide)
      if ((systems[i].flags & GBDPCG_SYS_FORWARD) && (systems[i].flags & GBDPCG_SYS_ADJOINT)) L.pair = true;
    L.pair = L.pair && pair28(dtype, L.nt, pair_only);
    launches.push_back(L);
    members.push_back(wide);
  }
  if (!lds_sys.empty()) {
    BatchLaunch L;
    L.nt = kThreads;
    launches.push_back(L);
    members.push_back(lds_sys);
  }

  std::vector<uint32_t> G(count, 0), items;
  std::vector<uint8_t> fused(count, 0);
  for (size_t li = 0; li < launches.size(); ++li) {
    const int rc = plan_launch(dtype, systems, members[li], cfg, prop, budget, launches[li], G, fused, items);
    if (rc < 0) return rc;
  }

  // device state
  auto plan = new gbdpcg_batch();
  plan->dtype = dtype;
  plan->device = dev;
  plan->count = count;
  plan->sys.assign(systems, systems + count);
  plan->wgs = G;
  plan->launches = launches;
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
              exchange_flags(cfg);
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
  // summary over the launches: widest grid, rounds one after the other
  gbdpcg_batch_info &in = plan->info;
  in.wgs_per_cu = ~0u;
  for (const BatchLaunch &L : launches) {
    in.grid = std::max(in.grid, L.grid);
    in.rounds += L.rounds;
    in.wgs_per_cu = std::min(in.wgs_per_cu, L.occ);
    in.lds_bytes = std::max(in.lds_bytes, (uint32_t)L.lds);
    in.total_items += L.items;
    in.est_cost += L.makespan;
  }
  in.systems = count;
  in.launches = (uint32_t)launches.size();
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
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (const BatchLaunch &L : plan->launches) {
    if (L.reg) {
      HIPCHK(reg_launch(plan->dtype, L.kn, L.nt, dummy, plan->d_sys, plan->d_items + L.item_off, L.rounds, 1, L.grid,
                        L.lds, st, RegVariant{true, L.bc, L.pair}));
    } else {
      HIPCHK(pcg_launch(plan->dtype, dummy, plan->d_sys, plan->d_items + L.item_off, L.rounds, 1, L.grid, L.lds,
                        st));
    }
  }
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
