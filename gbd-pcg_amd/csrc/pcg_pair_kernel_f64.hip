// Paired forward + adjoint register-resident kernel, double instances (the kernel
// is in pcg_reg_impl.h): both register layouts, stateSize 2, 6, 12, 14 (and
// mixed sizes at 512 threads), 256 and 512 threads (whole rows: 256 only);
// stateSize 28 at 256 (one lane per row) and 512 (S in LDS, paired items only).
#include "pcg_reg_impl.h"

namespace gbdpcg {

namespace {
using T = double;
template <int NS, int NT, bool BC>
const void *pp() {
  return reinterpret_cast<const void *>(&pcg_pair_kernel<T, NS, NT, BC>);
}
template <int NT>
const void *pp_n(int n, bool bc) {
  constexpr bool f64 = sizeof(T) == 8;
  if (bc) {
    if constexpr (f64 && NT != 256) {
      return nullptr;
    } else {
      switch (n) {
        case 12: return pp<12, NT, true>();
        case 14: return pp<14, NT, true>();
        default: return nullptr;
      }
    }
  }
  switch (n) {
    case 0:
    case -1:  // mixed sizes (paired plans hold rows of n <= 14 only)
      if constexpr (NT == 512) return pp<0, NT, false>();
      return nullptr;
    case 2: return pp<2, NT, false>();
    case 6: return pp<6, NT, false>();
    case 12: return pp<12, NT, false>();
    case 14: return pp<14, NT, false>();
    case 28:  // one lane per row at 256 threads; S in LDS at 512 (paired items only)
      return pp<28, NT, false>();
    default: return nullptr;
  }
}
}  // namespace

const void *pair_kernel_ptr_f64(int n, int nt, bool bc) {
  return nt == 256 ? pp_n<256>(n, bc) : nt == 512 ? pp_n<512>(n, bc) : nullptr;
}

}  // namespace gbdpcg
