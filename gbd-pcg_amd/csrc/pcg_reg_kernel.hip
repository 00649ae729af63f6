// Register-resident PCG kernel: fused-recurrence instances and the host side
// (kernel lookup, LDS size, occupancy, launch).  The kernel itself is in
// pcg_reg_impl.h; the reference-recurrence instances are compiled in
// pcg_reg_ref_kernel.hip.
#include "pcg_reg_impl.h"

namespace gbdpcg {

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

namespace {
template <typename T, int NS, int NT>
const void *reg_ptr() {
  return reinterpret_cast<const void *>(&pcg_reg_kernel<T, NS, NT, true>);
}

template <typename T, int NT>
const void *reg_ptr_n(int n) {
  switch (n) {
    case 2: return reg_ptr<T, 2, NT>();
    case 6: return reg_ptr<T, 6, NT>();
    case 12: return reg_ptr<T, 12, NT>();
    case 14: return reg_ptr<T, 14, NT>();
    case 28: return reg_ptr<T, 28, NT>();
    case 0:  // mixed sizes (batches), every specialised size: 512 threads only
      if constexpr (NT == 512) return reg_ptr<T, 0, NT>();
      return nullptr;
    case -1:  // mixed sizes up to 14 (batches whose n = 28 rows run apart): 512 only
      if constexpr (NT == 512) return reg_ptr<T, -1, NT>();
      return nullptr;
    default: return nullptr;
  }
}

const void *reg_kernel(int dtype, int n, int nt, RegVariant v) {
  if (v.pair) {
    if (!v.fused) return nullptr;
    return dtype == GBDPCG_F64 ? pair_kernel_ptr_f64(n, nt, v.bc) : pair_kernel_ptr_f32(n, nt, v.bc);
  }
  if (v.bc) return rowcast_kernel_ptr(dtype, n, nt, v.fused);
  if (!v.fused) return reg_ref_kernel_ptr(dtype, n, nt);
  if (dtype == GBDPCG_F64)
    return nt == 512 ? reg_ptr_n<double, 512>(n) : nt == 256 ? reg_ptr_n<double, 256>(n) : nullptr;
  return nt == 512 ? reg_ptr_n<float, 512>(n) : nt == 256 ? reg_ptr_n<float, 256>(n) : nullptr;
}
}  // namespace

bool reg_supported(int n, bool bc) {
  if (bc) return n == 12 || n == 14;
  return n == 2 || n == 6 || n == 12 || n == 14 || n == 28;
}

// LDS of one workgroup: ghost-extended vectors (5, paired: 8), the per-wave
// partials (3 per wave, paired: 6), the totals and the control words.
size_t reg_lds_bytes(int dtype, uint32_t n, uint32_t K, int nt, RegVariant v) {
  const size_t ts = dtype == GBDPCG_F64 ? 8 : 4;
  const uint32_t lpr = reg_lpr(n, nt, (int)ts);
  // paired instances that keep their S chunks in LDS: pair_s_slots chunks ahead
  // of the vectors, whose read-ahead is then 8 values
  const bool slds = v.pair && pair_s_lds(ts, nt, v.bc, (int)n);
  const size_t VL = (size_t)(K + 4) * padded_n(n) + (v.bc ? 16u : slds ? 8u : lpr * reg_chunk(n, lpr));
  // paired instances that keep their Pinv chunks in LDS: nt chunks ahead of the vectors
  const size_t pinv = v.pair && pair_pinv_lds(ts, nt, v.bc, (int)n) ? (size_t)nt * reg_chunk(n, lpr) : 0;
  const size_t smat = slds ? (size_t)pair_s_slots(nt, (int)n) * reg_chunk(n, lpr) : 0;
  if (v.pair) return (pinv + smat + 8 * VL + 6 * (size_t)(nt / 64) + 8 + 4) * ts;
  return (5 * VL + 3 * (size_t)(nt / 64) + 4 + 4) * ts;
}

// Largest knots per workgroup the lane layout allows: reg_lpr(n) lanes per row
// of the owned knots plus one ghost knot each side, and 4n ghost lanes (row
// broadcast: 16 lanes per knot).
uint32_t reg_max_knots(int dtype, uint32_t n, int nt, bool bc) {
  if (bc) return n <= 16 && (uint32_t)nt >= 4 * n ? (uint32_t)nt / 16 - 2 : 0;
  const uint32_t rows = (uint32_t)nt / reg_lpr(n, nt, dtype == GBDPCG_F64 ? 8 : 4);
  if (rows < 4 * n) return 0;
  return rows / n - 2;
}

hipError_t reg_prepare(int dtype, int n, int nt, size_t lds, RegVariant v) {
  const void *k = reg_kernel(dtype, n, nt, v);
  if (!k) return hipErrorInvalidValue;
  return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t reg_occupancy(int dtype, int n, int nt, size_t lds, int *wgs_per_cu, RegVariant v) {
  const void *k = reg_kernel(dtype, n, nt, v);
  if (!k) return hipErrorInvalidValue;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(wgs_per_cu, k, nt, lds);
}

hipError_t reg_launch(int dtype, int n, int nt, const KSys &one, const KSys *systems, const uint32_t *items,
                      uint32_t rounds, uint32_t stride, uint32_t grid, size_t lds, hipStream_t stream,
                      RegVariant v) {
  const void *k = reg_kernel(dtype, n, nt, v);
  if (!k) return hipErrorInvalidValue;
  KSys a0 = one;
  const KSys *a1 = systems;
  const uint32_t *a2 = items;
  uint32_t a3 = rounds, a4 = stride;
  void *args[] = {&a0, &a1, &a2, &a3, &a4};
  return hipLaunchKernel(k, dim3(grid), dim3(nt), args, lds, stream);
}

}  // namespace gbdpcg
