// Register-resident PCG kernel, reference-recurrence instances (single solves:
// 256 threads, the specialised state sizes).  The kernel is in pcg_reg_impl.h.
#include "pcg_reg_impl.h"

namespace gbdpcg {

namespace {
template <typename T, int NS>
const void *ref_ptr() {
  return reinterpret_cast<const void *>(&pcg_reg_kernel<T, NS, 256, false>);
}

template <typename T>
const void *ref_ptr_n(int n) {
  switch (n) {
    case 2: return ref_ptr<T, 2>();
    case 6: return ref_ptr<T, 6>();
    case 12: return ref_ptr<T, 12>();
    case 14: return ref_ptr<T, 14>();
    case 28: return ref_ptr<T, 28>();
    default: return nullptr;
  }
}
}  // namespace

const void *reg_ref_kernel_ptr(int dtype, int n, int nt) {
  if (nt != 256) return nullptr;
  return dtype == GBDPCG_F64 ? ref_ptr_n<double>(n) : ref_ptr_n<float>(n);
}

}  // namespace gbdpcg
