// Device helpers shared by the PCG kernels: exchange words, global-pointer
// casts and DPP-based cross-lane reductions (gfx950 / GFX9 DPP).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace gbdpcg {

// ---------------------------------------------------------------------------
// exchange words
// ---------------------------------------------------------------------------

// Global-address-space view of a generic pointer: accesses compile to global_*
// instead of flat_* (flat operations also count against the LDS wait counter).
template <typename T>
using gptr_t = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr_t<T> gptr(T *p) {
  return (gptr_t<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr_t<const T> gptr(const T *p) {
  return (gptr_t<const T>)p;
}

__device__ __forceinline__ void xstore(uint64_t *p, uint32_t payload, uint32_t epoch) {
  const uint64_t w = ((uint64_t)epoch << 32) | (uint64_t)payload;
  __hip_atomic_store(gptr(p), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Same word with a workgroup-scope store: it completes in the XCD's shared L2
// (the CU's L1 is write-through) instead of being written through to memory, and
// the device-scope polls below, which skip the L1, find it there.  Only valid
// when the writer and every reader of the word run on the same XCD; the kernels
// check that at run time (xcc_id) before using it.
__device__ __forceinline__ void xstore_xcd(uint64_t *p, uint32_t payload, uint32_t epoch) {
  const uint64_t w = ((uint64_t)epoch << 32) | (uint64_t)payload;
  __hip_atomic_store(gptr(p), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void xstore_sel(uint64_t *p, uint32_t payload, uint32_t epoch, bool xcd) {
  if (xcd)
    xstore_xcd(p, payload, epoch);
  else
    xstore(p, payload, epoch);
}

// XCD (accelerator complex die) this wave runs on: hardware register XCC_ID,
// bits [3:0].
__device__ __forceinline__ uint32_t xcc_id() { return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xFu; }

// A value every lane holds identically, moved to a scalar register: later
// arithmetic on it can use SGPR operands and branches on it become scalar.
__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ uint64_t xload(uint64_t *p) {
  return __hip_atomic_load(gptr(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Spins until *p carries `epoch`; returns false when the spin bound is reached.
__device__ __forceinline__ bool xwait(uint64_t *p, uint32_t epoch, uint32_t &payload) {
  for (uint32_t spin = 0; spin < kSpinLimit; ++spin) {
    const uint64_t w = xload(p);
    if ((uint32_t)(w >> 32) == epoch) {
      payload = (uint32_t)w;
      return true;
    }
  }
  return false;
}

// Fused multiply-add in the precision of T (__builtin_fma on floats would
// promote to double).
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Single-word wait for words that usually land before the caller looks (edge
// values): polls with a short sleep to keep the polling traffic low.
__device__ __forceinline__ bool xwait_lazy(uint64_t *p, uint32_t epoch, uint32_t &payload) {
  for (uint32_t spin = 0; spin < kSpinLimit; ++spin) {
    const uint64_t w = xload(p);
    if ((uint32_t)(w >> 32) == epoch) {
      payload = (uint32_t)w;
      return true;
    }
    __builtin_amdgcn_s_sleep(4);
  }
  return false;
}

// Waits until the NW consecutive words at p all carry `epoch` (one group of
// loads per poll).
template <int NW>
__device__ __forceinline__ bool xwait_words(uint64_t *p, uint32_t epoch, uint32_t (&out)[NW]) {
  for (uint32_t spin = 0; spin < kSpinLimit; ++spin) {
    uint64_t A[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) A[i] = xload(p + i);
    bool ok = true;
#pragma unroll
    for (int i = 0; i < NW; ++i) ok = ok && (uint32_t)(A[i] >> 32) == epoch;
    if (ok) {
#pragma unroll
      for (int i = 0; i < NW; ++i) out[i] = (uint32_t)A[i];
      return true;
    }
  }
  return false;
}


// a / b for the scalar recurrences.  FP32: hardware reciprocal (1 ulp) and one
// residual correction of the quotient, about 1 ulp in all (no denormal-mode
// switches, unlike the IEEE sequence, and four dependent operations instead of
// ten).  Meaningful for finite non-zero b only: callers test b and discard the
// quotient otherwise.
template <typename T>
__device__ __forceinline__ T qdiv(T a, T b) {
  if constexpr (sizeof(T) == 4) {
    const float r = __builtin_amdgcn_rcpf(b);
    const float q = a * r;
    const float res = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(res, r, q);
  } else {
    return a / b;
  }
}

template <typename T>
__device__ __forceinline__ uint32_t payload_word(T v, int q) {
  if constexpr (sizeof(T) == 4) {
    return __float_as_uint(v);
  } else {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return q == 0 ? (uint32_t)b : (uint32_t)(b >> 32);
  }
}

// ---------------------------------------------------------------------------
// reductions (deterministic: the same data give bit-identical sums in every
// lane, wave and workgroup)
// ---------------------------------------------------------------------------

// Cross-lane moves use DPP (a VALU operand modifier, a few cycles) rather than
// __shfl_xor, which lowers to ds_bpermute and costs an LDS round trip per step.
// GFX9 DPP controls:
constexpr int kDppQuadXor1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

// Value of `v` from the lane selected by CTRL; 0 where the source lane does not
// exist or the row is masked off.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

__device__ __forceinline__ float read_lane63(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double read_lane63(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Sum over the 64 lanes of a wave (all lanes active): inclusive scan inside each
// row of 16 lanes, then the row totals are folded into lane 63, which is
// broadcast.  A fixed tree: identical inputs give identical bits in every wave.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp<kDppRowShr1>(v);
  v += dpp<kDppRowShr2>(v);
  v += dpp<kDppRowShr4>(v);
  v += dpp<kDppRowShr8>(v);
  v += dpp<kDppRowBcast15, 0xa>(v);
  v += dpp<kDppRowBcast31, 0xc>(v);
  return read_lane63(v);
}

// Three independent wave sums with their DPP steps interleaved (a DPP read of a
// value written by the previous VALU op must wait; three chains hide that).
template <typename T>
__device__ __forceinline__ void wave_sum3(T &a, T &b, T &c) {
  a += dpp<kDppRowShr1>(a);
  b += dpp<kDppRowShr1>(b);
  c += dpp<kDppRowShr1>(c);
  a += dpp<kDppRowShr2>(a);
  b += dpp<kDppRowShr2>(b);
  c += dpp<kDppRowShr2>(c);
  a += dpp<kDppRowShr4>(a);
  b += dpp<kDppRowShr4>(b);
  c += dpp<kDppRowShr4>(c);
  a += dpp<kDppRowShr8>(a);
  b += dpp<kDppRowShr8>(b);
  c += dpp<kDppRowShr8>(c);
  a += dpp<kDppRowBcast15, 0xa>(a);
  b += dpp<kDppRowBcast15, 0xa>(b);
  c += dpp<kDppRowBcast15, 0xa>(c);
  a += dpp<kDppRowBcast31, 0xc>(a);
  b += dpp<kDppRowBcast31, 0xc>(b);
  c += dpp<kDppRowBcast31, 0xc>(c);
  a = read_lane63(a);
  b = read_lane63(b);
  c = read_lane63(c);
}

// Sum over each aligned group of 4 lanes; all 4 lanes get the same bits.
template <typename T>
__device__ __forceinline__ T quad_sum(T v) {
  v += dpp<kDppQuadXor2>(v);
  v += dpp<kDppQuadXor1>(v);
  return v;
}

}  // namespace gbdpcg
