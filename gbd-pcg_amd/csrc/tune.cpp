// Autotuner support: the choice cache (per process, optionally backed by a text
// file named by GBDPCG_TUNE_CACHE) and the batch tuner.  The single-solve tuner
// lives in api.cpp next to the layouts it enumerates.
//
// Batch tuning is a measured search over plans of the same systems: the cap on
// knots per workgroup (fewer, fuller workgroups per system vs more, emptier ones:
// exchange latency against per-iteration work and co-residency) and the threads
// per workgroup of the register kernel.  Each candidate plan runs on the systems'
// own data; lambda is restored before every run so all candidates solve the same
// problems, and at the end.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"

namespace gbdpcg {
void set_error(const char *fmt, ...);

namespace {
std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune;
bool g_file_loaded = false;

const char *cache_path() {
  const char *p = getenv("GBDPCG_TUNE_CACHE");
  return (p && *p) ? p : nullptr;
}

void load_file_locked() {
  if (g_file_loaded) return;
  g_file_loaded = true;
  const char *path = cache_path();
  if (!path) return;
  FILE *f = fopen(path, "r");
  if (!f) return;
  char key[512];
  unsigned a, b, c, d;
  while (fscanf(f, "%511s %u %u %u %u", key, &a, &b, &c, &d) == 5) g_tune[key] = TuneChoice{a, b, c, d};
  fclose(f);
}
}  // namespace

std::string device_tag(int dev) {
  // per device, once: reading the device properties costs microseconds and the
  // tag is needed on every automatic single solve
  static std::mutex mu;
  static std::map<int, std::string> tags;
  std::lock_guard<std::mutex> lk(mu);
  auto it = tags.find(dev);
  if (it != tags.end()) return it->second;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
  std::string arch = prop.gcnArchName;
  for (char &ch : arch)
    if (ch == ' ' || ch == ':') ch = '-';
  return tags[dev] = arch + "-cu" + std::to_string(prop.multiProcessorCount);
}

// Bumped whenever a choice is stored or the choices are dropped, so that callers
// holding a layout derived from the cache know to look again.
static std::atomic<uint64_t> g_tune_gen{1};
uint64_t tune_generation() { return g_tune_gen.load(std::memory_order_acquire); }

bool tune_lookup(const std::string &key, TuneChoice *out) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  load_file_locked();
  auto it = g_tune.find(key);
  if (it == g_tune.end()) return false;
  *out = it->second;
  return true;
}

void tune_store(const std::string &key, const TuneChoice &c) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  load_file_locked();
  g_tune[key] = c;
  g_tune_gen.fetch_add(1, std::memory_order_acq_rel);
  const char *path = cache_path();
  if (!path) return;
  FILE *f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %u %u %u %u\n", key.c_str(), c.a, c.b, c.c, c.d);
  fclose(f);
}

void tune_forget_all() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  g_file_loaded = true;  // the file (if any) is not re-read
  g_tune_gen.fetch_add(1, std::memory_order_acq_rel);
}

}  // namespace gbdpcg

using namespace gbdpcg;

#define HIPCHK(expr)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return GBDPCG_ERR_DEVICE;                                                              \
    }                                                                                        \
  } while (0)

namespace {

size_t tsz(int dtype) { return dtype == GBDPCG_F64 ? 8 : 4; }

// FNV-1a over the shapes and flags of the systems (order matters: the plan does).
uint64_t systems_hash(const gbdpcg_system_desc *s, uint32_t count) {
  uint64_t h = 1469598103934665603ULL;
  auto mix = [&](uint64_t v) {
    for (int i = 0; i < 8; ++i) {
      h ^= (v >> (8 * i)) & 0xFF;
      h *= 1099511628211ULL;
    }
  };
  for (uint32_t i = 0; i < count; ++i) {
    mix(s[i].n);
    mix(s[i].N);
    mix(s[i].flags);
    mix(s[i].wgs);
  }
  return h;
}

// Saved lambdas of the forward systems (restored before every run).
struct LambdaSnapshot {
  void *buf = nullptr;
  std::vector<std::pair<void *, size_t>> parts;  // (lambda, bytes)
  std::vector<size_t> off;
  ~LambdaSnapshot() {
    if (buf) (void)hipFree(buf);
  }
  int save(int dtype, const gbdpcg_system_desc *s, uint32_t count, hipStream_t st) {
    size_t total = 0;
    for (uint32_t i = 0; i < count; ++i)
      if (s[i].flags & GBDPCG_SYS_FORWARD) {
        const size_t b = (size_t)s[i].n * s[i].N * tsz(dtype);
        parts.push_back({s[i].lambda, b});
        off.push_back(total);
        total += (b + 255) & ~(size_t)255;
      }
    if (total == 0) return GBDPCG_OK;
    HIPCHK(hipMalloc(&buf, total));
    for (size_t i = 0; i < parts.size(); ++i)
      HIPCHK(hipMemcpyAsync(static_cast<char *>(buf) + off[i], parts[i].first, parts[i].second,
                            hipMemcpyDeviceToDevice, st));
    return GBDPCG_OK;
  }
  int restore(hipStream_t st) {
    for (size_t i = 0; i < parts.size(); ++i)
      HIPCHK(hipMemcpyAsync(parts[i].first, static_cast<char *>(buf) + off[i], parts[i].second,
                            hipMemcpyDeviceToDevice, st));
    return GBDPCG_OK;
  }
};

// Mean time of `trials` runs of a plan (lambda restored before each), after one
// untimed run.  Returns < 0 on error (incl. a run that timed out).
int time_plan(gbdpcg_batch *p, uint32_t count, LambdaSnapshot &snap, uint32_t trials, hipStream_t st, double *us) {
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  double total = 0.0;
  int rc = GBDPCG_OK;
  for (uint32_t t = 0; t <= trials && rc == GBDPCG_OK; ++t) {
    rc = snap.restore(st);
    if (rc < 0) break;
    if (hipEventRecord(a, st) != hipSuccess) rc = GBDPCG_ERR_DEVICE;
    if (rc == GBDPCG_OK) rc = gbdpcg_batch_run(p, st);
    if (rc == GBDPCG_OK && hipEventRecord(b, st) != hipSuccess) rc = GBDPCG_ERR_DEVICE;
    if (rc == GBDPCG_OK && hipEventSynchronize(b) != hipSuccess) rc = GBDPCG_ERR_DEVICE;
    float ms = 0.f;
    if (rc == GBDPCG_OK && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = GBDPCG_ERR_DEVICE;
    if (t > 0) total += ms * 1000.0;
  }
  if (rc == GBDPCG_OK) {
    std::vector<gbdpcg_result> res(count);
    rc = gbdpcg_batch_results(p, res.data(), st);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc < 0) return rc;
  *us = total / std::max<uint32_t>(trials, 1);
  return GBDPCG_OK;
}

}  // namespace

extern "C" int gbdpcg_batch_autotune(gbdpcg_batch **out, int dtype, const gbdpcg_system_desc *systems,
                                     uint32_t count, const gbdpcg_config *cfg, uint32_t trials,
                                     gbdpcg_tune_report *report) {
  if (!out || !systems || count == 0 || !cfg) {
    set_error("gbdpcg_batch_autotune: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  *out = nullptr;
  if (trials == 0) trials = 3;
  hipStream_t st = static_cast<hipStream_t>(cfg->stream);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
    set_error("gbdpcg_batch_autotune: cannot tune inside graph capture");
    return GBDPCG_ERR_INVALID;
  }
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  gbdpcg_tune_report rep{};
  const std::string key = "batch:" + device_tag(dev) + ":" + std::to_string(dtype) + ":" +
                          std::to_string(cfg->recurrence) + ":" + std::to_string(cfg->exit_mode) + ":" +
                          std::to_string(cfg->kernel) + ":" + std::to_string(cfg->threads) + ":" +
                          std::to_string(cfg->batch_split) + ":" + std::to_string(count) + ":" +
                          std::to_string(systems_hash(systems, count));
  TuneChoice hit;
  if (tune_lookup(key, &hit)) {
    gbdpcg_config c = *cfg;
    c.knots_per_wg = hit.a;
    c.threads = hit.b;
    c.kernel = hit.c;
    c.batch_split = hit.d;
    int rc = gbdpcg_batch_create(out, dtype, systems, count, &c, 0);
    if (rc == GBDPCG_OK) {
      gbdpcg_batch_info info{};
      gbdpcg_batch_info_get(*out, &info);
      rep.from_cache = 1;
      rep.knots_per_wg = hit.a;
      rep.threads = (uint32_t)batch_threads(*out);
      rep.wgs = info.grid;
      rep.kernel = (uint32_t)batch_kernel(*out);
      if (report) *report = rep;
      return GBDPCG_OK;
    }
    // the cached choice no longer plans (other device limits): tune again
  }

  // candidate knobs: kernel x threads per workgroup x one launch or one per
  // kernel class x knots-per-workgroup caps (identical plans are timed once, see
  // the deduplication below)
  uint32_t maxN = 0;
  for (uint32_t i = 0; i < count; ++i) maxN = std::max(maxN, systems[i].N);
  const uint32_t caps[] = {0, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64};
  // what the untuned plan would use
  const uint32_t def_nt = (cfg->threads == 256 || (cfg->threads == 0 && cfg->block[0] == 256)) ? 256u : 512u;
  const std::vector<uint32_t> nts = {def_nt, def_nt == 512 ? 256u : 512u};

  LambdaSnapshot snap;
  int rc = snap.save(dtype, systems, count, st);
  if (rc < 0) return rc;
  HIPCHK(hipStreamSynchronize(st));

  // the untuned plan's kernel first (its cap-0 plan is the default), then the
  // row-broadcast layout when the kernel is automatic
  std::vector<uint32_t> kernels = {cfg->kernel};
  if (cfg->kernel == GBDPCG_KERNEL_AUTO) kernels.push_back(GBDPCG_KERNEL_ROWCAST);

  // the requested launch split first (its cap-0 plan at the default threads is the default)
  const std::vector<uint32_t> splits = {cfg->batch_split, cfg->batch_split ? 0u : 1u};

  gbdpcg_batch *best = nullptr;
  double best_us = 1e300, default_us = -1.0;
  uint32_t best_cap = 0, best_nt = 512, best_kernel = cfg->kernel, best_split = cfg->batch_split;
  std::vector<std::vector<uint32_t>> seen;  // per-system workgroup vectors already timed
  std::vector<int> seen_nt, seen_k, seen_l;  // ... and the threads per workgroup, kernel and launches they ran with
  for (uint32_t kid : kernels)
  for (uint32_t nt : nts)
  for (uint32_t sp : splits) {
    for (uint32_t cap : caps) {
      if (cap > maxN) continue;
      gbdpcg_config c = *cfg;
      c.knots_per_wg = cap;
      c.threads = nt;
      c.kernel = kid;
      c.batch_split = sp;
      gbdpcg_batch *p = nullptr;
      if (gbdpcg_batch_create(&p, dtype, systems, count, &c, 0) != GBDPCG_OK) continue;  // does not plan
      std::vector<uint32_t> w(count);
      gbdpcg_batch_wgs(p, w.data());
      gbdpcg_batch_info pinfo{};
      gbdpcg_batch_info_get(p, &pinfo);
      bool dup = false;
      for (size_t j = 0; j < seen.size(); ++j)
        if (seen[j] == w && seen_nt[j] == batch_threads(p) && seen_k[j] == batch_kernel(p) &&
            seen_l[j] == (int)pinfo.launches)
          dup = true;
      if (dup) {
        gbdpcg_batch_destroy(p);
        continue;
      }
      seen.push_back(w);
      seen_nt.push_back(batch_threads(p));
      seen_k.push_back(batch_kernel(p));
      seen_l.push_back((int)pinfo.launches);
      double us = 0.0;
      rc = time_plan(p, count, snap, trials, st, &us);
      if (rc < 0) {
        gbdpcg_batch_destroy(p);
        if (best) gbdpcg_batch_destroy(best);
        snap.restore(st);
        (void)hipStreamSynchronize(st);
        return rc;
      }
      ++rep.candidates;
      if (cap == 0 && nt == def_nt && kid == cfg->kernel && sp == cfg->batch_split) default_us = us;
      if (us < best_us) {
        if (best) gbdpcg_batch_destroy(best);
        best = p;
        best_us = us;
        best_cap = cap;
        best_nt = nt;
        best_kernel = kid;
        best_split = sp;
      } else {
        gbdpcg_batch_destroy(p);
      }
    }
  }
  rc = snap.restore(st);
  if (rc == GBDPCG_OK && hipStreamSynchronize(st) != hipSuccess) rc = GBDPCG_ERR_DEVICE;
  if (!best) {
    set_error("gbdpcg_batch_autotune: no candidate plan fits these systems");
    return GBDPCG_ERR_TOO_LARGE;
  }
  if (rc < 0) {
    gbdpcg_batch_destroy(best);
    return rc;
  }
  gbdpcg_batch_info info{};
  gbdpcg_batch_info_get(best, &info);
  rep.knots_per_wg = best_cap;
  rep.threads = (uint32_t)batch_threads(best);
  rep.wgs = info.grid;
  rep.kernel = (uint32_t)batch_kernel(best);
  rep.best_us = best_us;
  rep.default_us = default_us;
  tune_store(key, TuneChoice{best_cap, best_nt, best_kernel, best_split});
  if (report) *report = rep;
  *out = best;
  return GBDPCG_OK;
}
