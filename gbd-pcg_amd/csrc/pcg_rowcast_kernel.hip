// Register-resident PCG kernel, row-broadcast layout (stateSize 12 and 14): one
// lane per matrix row and 16 lanes per knot, the vector window of every row
// broadcast within its 16-lane row by DPP.  The kernel is in pcg_reg_impl.h.
#include "pcg_reg_impl.h"

namespace gbdpcg {

namespace {
template <typename T, int NS, int NT, bool FUSED>
const void *rc_ptr() {
  return reinterpret_cast<const void *>(&pcg_rowcast_kernel<T, NS, NT, FUSED>);
}

template <typename T, int NT, bool FUSED>
const void *rc_ptr_n(int n) {
  switch (n) {
    case 12: return rc_ptr<T, 12, NT, FUSED>();
    case 14: return rc_ptr<T, 14, NT, FUSED>();
    default: return nullptr;
  }
}
}  // namespace

// Fused recurrence: 64, 256 and 512 threads; reference recurrence: 64 and 256.
// 64 threads: one wave per rank (two knots), no cross-wave reductions.
const void *rowcast_kernel_ptr(int dtype, int n, int nt, bool fused) {
  const bool f64 = dtype == GBDPCG_F64;
  if (nt == 64) return fused ? (f64 ? rc_ptr_n<double, 64, true>(n) : rc_ptr_n<float, 64, true>(n))
                             : (f64 ? rc_ptr_n<double, 64, false>(n) : rc_ptr_n<float, 64, false>(n));
  if (!fused) {
    if (nt != 256) return nullptr;
    return f64 ? rc_ptr_n<double, 256, false>(n) : rc_ptr_n<float, 256, false>(n);
  }
  if (nt == 256) return f64 ? rc_ptr_n<double, 256, true>(n) : rc_ptr_n<float, 256, true>(n);
  if (nt == 512) return f64 ? rc_ptr_n<double, 512, true>(n) : rc_ptr_n<float, 512, true>(n);
  return nullptr;
}

}  // namespace gbdpcg
