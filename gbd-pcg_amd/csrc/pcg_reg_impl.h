// Register-resident PCG kernels for gfx950: the latency path (templates,
// instantiated in pcg_reg_kernel.hip (fused recurrence, column chunks),
// pcg_reg_ref_kernel.hip (reference recurrence), pcg_rowcast_kernel.hip (row
// broadcast) and pcg_pair_kernel_f32/_f64.hip (forward + adjoint together)).
//
// Same method and exchange protocol as pcg_kernel.hip, with every block row a
// workgroup owns held in registers instead of LDS.  FUSED: the single-reduction
// (Chronopoulos-Gear) form, described below; reference form
// (include/pcg.cuh:118-208): two exchanges per iteration, the first carrying
// p'v and the edge knots of v = S p, the second r'z and r'r only (z = Pinv r is
// formed on one ghost knot each side from the resident ghost rows of Pinv, so
// p needs no halo).  Fused form:
//   * LPR lanes (reg_lpr: 2 for most sizes, 1 for n <= 2 and for the n = 28
//     layouts that hold whole rows) share one matrix row; each lane keeps a
//     CW-column chunk of that row of Pinv (rows of the owned knots plus one
//     ghost knot each side) and of S (owned knots) in VGPRs for the whole
//     solve, loaded from HBM once.  Row-broadcast layout (BC): one lane per row,
//     16 lanes per knot, whole rows, the vector window broadcast by DPP;
//   * per iteration: one pass u = Pinv r (owned + ghost rows), one pass w = S u
//     (owned rows), then one cross-workgroup exchange carrying (r'u, w'u, r'r)
//     and the two edge knots of w each side;
//   * the lanes that produce an edge value publish it themselves, and the
//     wave that polls the scalars also reduces them (DPP), so an iteration has
//     four workgroup barriers and one exchange.
// LDS holds only ghost-extended vectors (a few KB), so several workgroups fit a
// CU when the batched path uses this kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"
#include "internal.h"

namespace gbdpcg {

namespace {
// threadIdx.x through an empty asm: index arithmetic derived from it is then
// recomputed where it is used instead of being hoisted to the kernel entry and
// kept live (in VGPRs) across the whole solve.
__device__ __forceinline__ int opaque_tid() {
  int t = (int)threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

struct RegOut {
  int iters, status;
  double eta, rr;
};

// Sum over the waves of a workgroup of partial l (stride S per wave), as a
// pairwise tree in a fixed order: three dependent adds for eight waves instead
// of seven.
template <int NW, int S, typename T>
__device__ __forceinline__ T waves_total(const T *wsum, int l) {
  if constexpr (NW == 1) {
    return wsum[l];
  } else {
    return waves_total<NW / 2, S>(wsum, l) + waves_total<NW - NW / 2, S>(wsum + (NW / 2) * S, l);
  }
}

// Status of one exchange within the item (all threads agree after a gather).
// BC: row-broadcast layout (one lane per matrix row, 16 lanes per knot, every
// lane holding its whole row of 3n columns).
template <typename T, int NS, int NT, bool BC = false>
struct RegItem {
  static constexpr int n = NS;
  static constexpr int nP = (int)padded_n(NS);
  static constexpr int LPR = BC ? 1 : (int)reg_lpr(NS, NT, (int)sizeof(T));  // lanes per matrix row
  static constexpr int CW = BC ? 3 * NS : (int)reg_chunk(NS, LPR);           // columns per lane
  static constexpr int WPV = (int)(sizeof(T) / 4);
  static constexpr int RW = (int)rec_words(NS, WPV);
  static constexpr int NWAVES = NT / 64;

  // geometry
  int N, G, rank, a, b, K, VL;
  uint64_t *slots;
  uint32_t base, seq;
  double tol;
  bool rel;
  bool xcd;     // every rank on this workgroup's XCD: exchange stores stay in L2
  bool xcd_ok;  // the system allows XCD-local stores (KF_NOXCD clear)
  // lane roles
  int q, c, m, row, kg;
  bool act_u, act_w, owner, ghost;
  int gpos;
  // LDS
  T *rv, *uv, *wv, *xv, *lv, *wsum, *tot;
  int *ctl;

  __device__ __forceinline__ uint64_t *rec(int type, int g) const {
    return slots + (size_t)(type * G + g) * RW;
  }

  // Owner lanes publish their value into the edge slots of the current exchange
  // (value kRecScalars + e*n + row for own knot a+e, kRecScalars + 2n + e*n + row
  // for own knot b-1-e).  A lane has at most two such slots (K >= 2), slot_lo for
  // its left-edge role and slot_hi for its right-edge one, each -1 when absent and
  // tagged with its depth e in bit 16 (edge = 1 publishes depth 0 only).
  int slot_lo, slot_hi;

  __device__ __forceinline__ void set_edge_slots() {
    const int kl = m - 1;
    slot_lo = -1;
    slot_hi = -1;
    if (!owner) return;
    if (kl < 2) slot_lo = (kRecScalars + kl * n + row) | (kl << 16);
    const int e = K - 1 - kl;
    if (e < 2) slot_hi = (kRecScalars + 2 * n + e * n + row) | (e << 16);
  }

  // voff: 0 for the first edge block of a record, 4n for the second (paired
  // solves publish the edges of two vectors per exchange).
  template <bool XCD>
  __device__ __forceinline__ void store_edges(uint64_t *r, T val, int edge, uint32_t epoch, int voff) const {
    const int lim = edge << 16;
    if (slot_lo >= 0 && slot_lo < lim) {
#pragma unroll
      for (int qq = 0; qq < WPV; ++qq) {
        uint64_t *p = r + ((slot_lo & 0xFFFF) + voff) * WPV + qq;
        if constexpr (XCD) xstore_xcd(p, payload_word(val, qq), epoch); else xstore(p, payload_word(val, qq), epoch);
      }
    }
    if (slot_hi >= 0 && slot_hi < lim) {
#pragma unroll
      for (int qq = 0; qq < WPV; ++qq) {
        uint64_t *p = r + ((slot_hi & 0xFFFF) + voff) * WPV + qq;
        if constexpr (XCD) xstore_xcd(p, payload_word(val, qq), epoch); else xstore(p, payload_word(val, qq), epoch);
      }
    }
  }

  __device__ __forceinline__ void publish_edge(T val, int edge, int voff = 0) const {
    if (G == 1) return;
    const uint32_t epoch = base + seq;
    uint64_t *r = rec((int)(seq & 1u), rank);
    if (xcd)
      store_edges<true>(r, val, edge, epoch, voff);
    else
      store_edges<false>(r, val, edge, epoch, voff);
  }

  // Workgroup totals of three per-lane partials (DPP per wave, then the waves in
  // order by thread 0), published as this rank's three scalars; with one
  // workgroup they go straight to tot.  One barrier.  With `placement`, the
  // second and third scalars are replaced by x and x^2, x = 1 + this XCD's id,
  // so that after the gather every rank can tell (identically) whether all ranks
  // share one XCD: G * sum(x^2) == sum(x)^2 exactly when all x are equal.
  __device__ __forceinline__ void reduce_publish(T s0, T s1, T s2, bool placement = false) const {
    wave_sum3(s0, s1, s2);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      wsum[w * 3 + 0] = s0;
      wsum[w * 3 + 1] = s1;
      wsum[w * 3 + 2] = s2;
    }
    __syncthreads();
    // lanes 0..2 of wave 0: scalar l summed over the waves in order, published
    const int l = threadIdx.x;
    if (l >= 3) return;
    T tv = waves_total<NWAVES, 3>(wsum, l);
    if (G == 1) {
      tot[l] = tv;
      return;
    }
    if (placement && l > 0) {
      const T x = (T)(1 + xcc_id());
      tv = l == 1 ? x : x * x;
    }
    const uint32_t epoch = base + seq;
    uint64_t *r = rec((int)(seq & 1u), rank);
#pragma unroll
    for (int qq = 0; qq < WPV; ++qq) xstore_sel(r + l * WPV + qq, payload_word(tv, qq), epoch, xcd);
  }

  static __device__ __forceinline__ T word_value(const uint32_t *w, int v) {
    if constexpr (WPV == 1) {
      return __uint_as_float(w[v]);
    } else {
      return __longlong_as_double((long long)(((uint64_t)w[2 * v + 1] << 32) | w[2 * v]));
    }
  }

  // Source word and LDS destination of edge entry i (payload word i of the
  // neighbours' edge knots, left neighbour's first) of edge block voff.
  __device__ __forceinline__ void edge_slot(int i, int nleft, int type, int voff, T *vec, uint64_t *&src,
                                            uint32_t *&dst) const {
    if (i < nleft) {  // left neighbour's right edge: knot a-1-e = local 1-e
      const int vi = i / WPV, qq = i - vi * WPV;
      const int e = vi / n, rr = vi - e * n;
      src = rec(type, rank - 1) + (kRecScalars + voff + 2 * n + e * n + rr) * WPV + qq;
      dst = reinterpret_cast<uint32_t *>(vec + (1 - e) * nP + rr) + qq;
    } else {  // right neighbour's left edge: knot b+e = local K+2+e
      const int ii = i - nleft;
      const int vi = ii / WPV, qq = ii - vi * WPV;
      const int e = vi / n, rr = vi - e * n;
      src = rec(type, rank + 1) + (kRecScalars + voff + e * n + rr) * WPV + qq;
      dst = reinterpret_cast<uint32_t *>(vec + (K + 2 + e) * nP + rr) + qq;
    }
  }

  // One-wave workgroups: lane t polls rank t's three totals and edge entries t
  // and t + 64 in one spin loop (one round trip when everything has landed);
  // ranks beyond 64 are polled afterwards.  Sums as the multi-wave gather does.
  __device__ __forceinline__ bool gather_one_wave(T *vec, int edge, uint32_t epoch, int type, int t) {
    constexpr int NW = 3 * WPV;
    const int en = edge * n * WPV;
    const int nleft = rank > 0 ? en : 0;
    const int ne = nleft + (rank < G - 1 ? en : 0);  // <= 4 * 16 * 2 = 128
    uint64_t *src0 = nullptr, *src1 = nullptr;
    uint32_t *dst0 = nullptr, *dst1 = nullptr;
    if (t < ne) edge_slot(t, nleft, type, 0, vec, src0, dst0);
    if (t + 64 < ne) edge_slot(t + 64, nleft, type, 0, vec, src1, dst1);
    bool need_s = t < G, need0 = src0 != nullptr, need1 = src1 != nullptr;
    uint32_t wv_[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) wv_[i] = 0;
    uint32_t v0 = 0, v1 = 0;
    for (uint32_t spin = 0; spin < kSpinLimit && (need_s || need0 || need1); ++spin) {
      if (need_s) {
        uint64_t A[NW];
        uint64_t *p = rec(type, t);
#pragma unroll
        for (int i = 0; i < NW; ++i) A[i] = xload(p + i);
        bool all = true;
#pragma unroll
        for (int i = 0; i < NW; ++i) all = all && (uint32_t)(A[i] >> 32) == epoch;
        if (all) {
#pragma unroll
          for (int i = 0; i < NW; ++i) wv_[i] = (uint32_t)A[i];
          need_s = false;
        }
      }
      if (need0) {
        const uint64_t w = xload(src0);
        if ((uint32_t)(w >> 32) == epoch) {
          v0 = (uint32_t)w;
          need0 = false;
        }
      }
      if (need1) {
        const uint64_t w = xload(src1);
        if ((uint32_t)(w >> 32) == epoch) {
          v1 = (uint32_t)w;
          need1 = false;
        }
      }
    }
    bool ok = !(need_s || need0 || need1);
    if (src0 && !need0) *dst0 = v0;
    if (src1 && !need1) *dst1 = v1;
    T acc0 = 0, acc1 = 0, acc2 = 0;
    if (t < G && !need_s) {
      acc0 = word_value(wv_, 0);
      acc1 = word_value(wv_, 1);
      acc2 = word_value(wv_, 2);
    }
    for (int g = t + 64; g < G; g += 64) {
      uint32_t w2[NW];
      if (!xwait_words<NW>(rec(type, g), epoch, w2)) {
        ok = false;
        break;
      }
      acc0 += word_value(w2, 0);
      acc1 += word_value(w2, 1);
      acc2 += word_value(w2, 2);
    }
    wave_sum3(acc0, acc1, acc2);
    if (t == 0) {
      tot[0] = acc0;
      tot[1] = acc1;
      tot[2] = acc2;
    }
    return ok;
  }

  // Polls the current exchange.  Wave 0 gathers every rank's wave totals and
  // reduces them (waves in order, then ranks by a fixed DPP tree) into tot; the
  // other waves copy the neighbours' `edge` edge knots into the ghost knots of
  // `vec` (one-wave workgroups: gather_one_wave).  Ends with a barrier after
  // which t0..t2 hold the totals in every thread; returns false (uniformly) if
  // a wait hit its bound.
  __device__ __forceinline__ bool gather(T *vec, int edge, T &t0, T &t1, T &t2) {
    if (G == 1) {
      __syncthreads();
      t0 = uniform(tot[0]);
      t1 = uniform(tot[1]);
      t2 = uniform(tot[2]);
      ++seq;
      return true;
    }
    const uint32_t epoch = base + seq;
    const int type = (int)(seq & 1u);
    bool ok = true;
    const int t = opaque_tid();
    if constexpr (NT == 64) {
      ok = gather_one_wave(vec, edge, epoch, type, t);
    } else if (t < 64) {
      constexpr int NW = 3 * WPV;
      T acc0 = 0, acc1 = 0, acc2 = 0;
      for (int g = t; g < G; g += 64) {
        uint32_t wv_[NW];
        if (!xwait_words<NW>(rec(type, g), epoch, wv_)) {
          ok = false;
          break;
        }
        acc0 += word_value(wv_, 0);
        acc1 += word_value(wv_, 1);
        acc2 += word_value(wv_, 2);
      }
      wave_sum3(acc0, acc1, acc2);
      if (t == 0) {
        tot[0] = acc0;
        tot[1] = acc1;
        tot[2] = acc2;
      }
    } else {
      const int en = edge * n * WPV;
      const int nleft = rank > 0 ? en : 0;
      const int nright = rank < G - 1 ? en : 0;
      for (int i = t - 64; i < nleft + nright; i += NT - 64) {
        uint64_t *src;
        uint32_t *dst;
        edge_slot(i, nleft, type, 0, vec, src, dst);
        uint32_t val = 0;
        if (xcd ? xwait(src, epoch, val) : xwait_lazy(src, epoch, val)) {
          *dst = val;
        } else {
          ok = false;
          break;
        }
      }
    }
    if (!ok) ctl[0] = 1;
    __syncthreads();
    ++seq;
    t0 = uniform(tot[0]);
    t1 = uniform(tot[1]);
    t2 = uniform(tot[2]);
    return __builtin_amdgcn_readfirstlane(ctl[0]) == 0;
  }

  // Paired solves: six workgroup totals (the first three for vector A, the last
  // three for B; each scalar reduced exactly as reduce_publish reduces it),
  // wsum holding 6 per wave.  Lanes 0..5 publish.  placement as in
  // reduce_publish, for scalars 1 and 2.
  __device__ __forceinline__ void reduce_publish6(T s0, T s1, T s2, T s3, T s4, T s5,
                                                  bool placement = false) const {
    wave_sum3(s0, s1, s2);
    wave_sum3(s3, s4, s5);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      wsum[w * 6 + 0] = s0;
      wsum[w * 6 + 1] = s1;
      wsum[w * 6 + 2] = s2;
      wsum[w * 6 + 3] = s3;
      wsum[w * 6 + 4] = s4;
      wsum[w * 6 + 5] = s5;
    }
    __syncthreads();
    const int l = threadIdx.x;
    if (l >= 6) return;
    T tv = waves_total<NWAVES, 6>(wsum, l);
    if (G == 1) {
      tot[l] = tv;
      return;
    }
    if (placement && (l == 1 || l == 2)) {
      const T x = (T)(1 + xcc_id());
      tv = l == 1 ? x : x * x;
    }
    const uint32_t epoch = base + seq;
    uint64_t *r = rec((int)(seq & 1u), rank);
#pragma unroll
    for (int qq = 0; qq < WPV; ++qq) xstore_sel(r + l * WPV + qq, payload_word(tv, qq), epoch, xcd);
  }

  // gather for paired solves: six totals, and the neighbours' `edge` edge knots
  // of vA (first edge block) and vB (second) into their ghost knots.
  __device__ __forceinline__ bool gather_pair(T *vA, T *vB, int edge, T (&tt)[6]) {
    if (G == 1) {
      __syncthreads();
#pragma unroll
      for (int l = 0; l < 6; ++l) tt[l] = uniform(tot[l]);
      ++seq;
      return true;
    }
    const uint32_t epoch = base + seq;
    const int type = (int)(seq & 1u);
    bool ok = true;
    const int t = opaque_tid();
    if (t < 64) {
      constexpr int NW = 6 * WPV;
      T acc[6] = {0, 0, 0, 0, 0, 0};
      for (int g = t; g < G; g += 64) {
        uint32_t wv_[NW];
        if (!xwait_words<NW>(rec(type, g), epoch, wv_)) {
          ok = false;
          break;
        }
#pragma unroll
        for (int l = 0; l < 6; ++l) acc[l] += word_value(wv_, l);
      }
      wave_sum3(acc[0], acc[1], acc[2]);
      wave_sum3(acc[3], acc[4], acc[5]);
      if (t == 0) {
#pragma unroll
        for (int l = 0; l < 6; ++l) tot[l] = acc[l];
      }
    } else {
      const int en = edge * n * WPV;
      const int nleft = rank > 0 ? en : 0;
      const int nright = rank < G - 1 ? en : 0;
      const int per = nleft + nright;
      for (int i = t - 64; i < 2 * per; i += NT - 64) {
        const int v = i >= per ? 1 : 0;
        const int iv = i - v * per;
        uint64_t *src;
        uint32_t *dst;
        edge_slot(iv, nleft, type, v * 4 * n, v ? vB : vA, src, dst);
        uint32_t val = 0;
        if (xcd ? xwait(src, epoch, val) : xwait_lazy(src, epoch, val)) {
          *dst = val;
        } else {
          ok = false;
          break;
        }
      }
    }
    if (!ok) ctl[0] = 1;
    __syncthreads();
    ++seq;
#pragma unroll
    for (int l = 0; l < 6; ++l) tt[l] = uniform(tot[l]);
    return __builtin_amdgcn_readfirstlane(ctl[0]) == 0;
  }
};

// Four accumulator chains (the FMA latency, not the issue rate, bounds a short
// dot product), summed in a fixed order.
template <typename T, int CW>
__device__ __forceinline__ T chunk_dot(const T (&mat)[CW], const T *__restrict__ x) {
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
#pragma unroll
  for (int j = 0; j < CW; j += 4) {
    if constexpr (sizeof(T) == 4) {
      const float4 xv = *reinterpret_cast<const float4 *>(x + j);
      a0 = __builtin_fmaf(mat[j + 0], xv.x, a0);
      a1 = __builtin_fmaf(mat[j + 1], xv.y, a1);
      a2 = __builtin_fmaf(mat[j + 2], xv.z, a2);
      a3 = __builtin_fmaf(mat[j + 3], xv.w, a3);
    } else {
      const double2 x0 = *reinterpret_cast<const double2 *>(x + j);
      const double2 x1 = *reinterpret_cast<const double2 *>(x + j + 2);
      a0 = fmaT(mat[j + 0], x0.x, a0);
      a1 = fmaT(mat[j + 1], x0.y, a1);
      a2 = fmaT(mat[j + 2], x1.x, a2);
      a3 = fmaT(mat[j + 3], x1.y, a3);
    }
  }
  return (a0 + a1) + (a2 + a3);
}

// The paired kernels' FP32 instances held to 128 VGPRs keep their Pinv chunks in
// LDS rather than in registers (see pair_pinv_lds): float4 block b of lane t's
// chunk sits at float4 index b * NT + t, so a wave's 16-byte reads of one block
// are contiguous and conflict-free.  Two dot products (the two sides of a
// paired solve) share each matrix read; each side's four accumulator chains
// follow chunk_dot's order, so the bits are those of chunk_dot.
template <int CW, int NT>
__device__ __forceinline__ void chunk_dot2_lds(const float *__restrict__ plds, int tid, const float *__restrict__ xa,
                                               const float *__restrict__ xb, float &ya, float &yb) {
  float a0 = 0, a1 = 0, a2 = 0, a3 = 0, b0 = 0, b1 = 0, b2 = 0, b3 = 0;
#pragma unroll
  for (int j = 0; j < CW; j += 4) {
    const float4 mv = *reinterpret_cast<const float4 *>(plds + ((size_t)(j / 4) * NT + tid) * 4);
    const float4 u = *reinterpret_cast<const float4 *>(xa + j);
    const float4 v = *reinterpret_cast<const float4 *>(xb + j);
    a0 = __builtin_fmaf(mv.x, u.x, a0);
    a1 = __builtin_fmaf(mv.y, u.y, a1);
    a2 = __builtin_fmaf(mv.z, u.z, a2);
    a3 = __builtin_fmaf(mv.w, u.w, a3);
    b0 = __builtin_fmaf(mv.x, v.x, b0);
    b1 = __builtin_fmaf(mv.y, v.y, b1);
    b2 = __builtin_fmaf(mv.z, v.z, b2);
    b3 = __builtin_fmaf(mv.w, v.w, b3);
  }
  ya = (a0 + a1) + (a2 + a3);
  yb = (b0 + b1) + (b2 + b3);
}

// FP64 chunks held in LDS as 16-byte blocks: block b of slot i at double2 index
// b * SL + i (a wave's reads of one block are contiguous).  Two dot products
// (the two sides of a paired solve) share each matrix read, each in chunk_dot's
// order, so the bits are those of chunk_dot; chunk_dot_lds_d is one of them.
template <int CW, int SL>
__device__ __forceinline__ void chunk_dot2_lds_d(const double *__restrict__ mlds, int idx,
                                                 const double *__restrict__ xa, const double *__restrict__ xb,
                                                 double &ya, double &yb) {
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, b0 = 0, b1 = 0, b2 = 0, b3 = 0;
#pragma unroll
  for (int j = 0; j < CW; j += 4) {
    const double2 m0 = *reinterpret_cast<const double2 *>(mlds + ((size_t)(j / 2) * SL + idx) * 2);
    const double2 m1 = *reinterpret_cast<const double2 *>(mlds + ((size_t)(j / 2 + 1) * SL + idx) * 2);
    const double2 u0 = *reinterpret_cast<const double2 *>(xa + j);
    const double2 u1 = *reinterpret_cast<const double2 *>(xa + j + 2);
    const double2 v0 = *reinterpret_cast<const double2 *>(xb + j);
    const double2 v1 = *reinterpret_cast<const double2 *>(xb + j + 2);
    a0 = __builtin_fma(m0.x, u0.x, a0);
    a1 = __builtin_fma(m0.y, u0.y, a1);
    a2 = __builtin_fma(m1.x, u1.x, a2);
    a3 = __builtin_fma(m1.y, u1.y, a3);
    b0 = __builtin_fma(m0.x, v0.x, b0);
    b1 = __builtin_fma(m0.y, v0.y, b1);
    b2 = __builtin_fma(m1.x, v1.x, b2);
    b3 = __builtin_fma(m1.y, v1.y, b3);
  }
  ya = (a0 + a1) + (a2 + a3);
  yb = (b0 + b1) + (b2 + b3);
}

template <int CW, int SL>
__device__ __forceinline__ double chunk_dot_lds_d(const double *__restrict__ mlds, int idx,
                                                  const double *__restrict__ x) {
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
#pragma unroll
  for (int j = 0; j < CW; j += 4) {
    const double2 m0 = *reinterpret_cast<const double2 *>(mlds + ((size_t)(j / 2) * SL + idx) * 2);
    const double2 m1 = *reinterpret_cast<const double2 *>(mlds + ((size_t)(j / 2 + 1) * SL + idx) * 2);
    const double2 u0 = *reinterpret_cast<const double2 *>(x + j);
    const double2 u1 = *reinterpret_cast<const double2 *>(x + j + 2);
    a0 = __builtin_fma(m0.x, u0.x, a0);
    a1 = __builtin_fma(m0.y, u0.y, a1);
    a2 = __builtin_fma(m1.x, u1.x, a2);
    a3 = __builtin_fma(m1.y, u1.y, a3);
  }
  return (a0 + a1) + (a2 + a3);
}

// Sum over the LPR lanes sharing one matrix row; all of them get the same bits.
template <int LPR, typename T>
__device__ __forceinline__ T group_sum(T v) {
  if constexpr (LPR == 2) {
    v += dpp<kDppQuadXor1>(v);
  } else if constexpr (LPR == 4) {
    v = quad_sum(v);
  }
  return v;
}

// Loads this lane's chunk of row `row` of block row kg of a cbtd matrix
// (padded row-major columns [c*CW, (c+1)*CW) of the n x 3nP block row).
constexpr int kLoadTables = 0, kLoadArith = 1, kLoadGuarded = 2, kLoadRows = 3;

// Per-chunk column tables of a block row (lane chunk c of CW padded columns):
// the element offset of each column (0 for a padded column) and bit masks of
// the valid columns, of those in the first block (S_{k,k-1}) and in the last
// (S_{k,k+1}).
template <int NS, int CW>
struct ChunkTables {
  uint32_t off[2][CW];
  uint32_t valid[2], seg0[2], seg2[2];
};
template <int NS, int CW>
__host__ __device__ constexpr ChunkTables<NS, CW> chunk_tables() {
  constexpr int n = NS, nP = (int)padded_n(NS);
  ChunkTables<NS, CW> t{};
  for (int c = 0; c < 2; ++c) {
    t.valid[c] = t.seg0[c] = t.seg2[c] = 0u;
    for (int j = 0; j < CW; ++j) {
      const int col = c * CW + j, seg = col / nP, cc = col - seg * nP;
      const bool ok = seg < 3 && cc < n;
      t.off[c][j] = ok ? (uint32_t)((seg * n + cc) * n) : 0u;
      if (ok) t.valid[c] |= 1u << j;
      if (ok && seg == 0) t.seg0[c] |= 1u << j;
      if (ok && seg == 2) t.seg2[c] |= 1u << j;
    }
  }
  return t;
}

// Matrix chunk loads.  kLoadTables: branch-free, offsets from compile-time
// tables (fewest instructions, but the table constants take VGPRs: FP32
// 256-thread fixed-size instances, the single-solve layout); kLoadArith:
// branch-free, offsets computed per element (other fixed sizes n <= 14);
// kLoadGuarded: one guarded load per element (an exec-mask branch each, so the
// loads issue slowly), kept where the branch-free forms' live masks cost more
// than they save (the paired kernel's 128-VGPR allocation, the runtime-n
// instances); kLoadRows: whole rows (one lane per row, CW = 3 nP), whose guards
// are uniform over each block of n columns: one branch per block (the n = 28
// one-lane-per-row instances).
template <typename T, int NS, int CW, int MODE = kLoadTables>
__device__ __forceinline__ void load_chunk(T (&dst)[CW], const T *__restrict__ A, int kg, int row, int c, int N,
                                           bool active) {
  constexpr int n = NS, nP = (int)padded_n(NS);
  // opaque copies: keeps the per-lane offsets below from being hoisted out of the
  // batched item loop (they would stay live, in registers, for the whole solve)
  int cl = c, rl = row;
  asm volatile("" : "+v"(cl), "+v"(rl));
  if constexpr (MODE == kLoadRows) {
    static_assert(CW == 3 * nP, "whole-row loads");
    const bool rowok = active && kg >= 0 && kg < N;
#pragma unroll
    for (int s3 = 0; s3 < 3; ++s3) {
      if (rowok && !(kg == 0 && s3 == 0) && !(kg == N - 1 && s3 == 2)) {
#pragma unroll
        for (int cc = 0; cc < nP; ++cc)
          dst[s3 * nP + cc] = cc < n ? gptr(A)[(size_t)kg * 3 * n * n + (size_t)(s3 * n + cc) * n + rl] : (T)0;
      } else {
#pragma unroll
        for (int cc = 0; cc < nP; ++cc) dst[s3 * nP + cc] = (T)0;
      }
    }
  } else if constexpr (MODE == kLoadGuarded) {
#pragma unroll
    for (int j = 0; j < CW; ++j) {
      const int col = cl * CW + j;
      const int seg = col / nP, cc = col - seg * nP;
      T v = 0;
      if (active && kg >= 0 && kg < N && seg < 3 && cc < n && !(kg == 0 && seg == 0) && !(kg == N - 1 && seg == 2))
        v = gptr(A)[(size_t)kg * 3 * n * n + (size_t)(seg * n + cc) * n + rl];
      dst[j] = v;
    }
  } else if constexpr (MODE == kLoadArith) {
    // branch-free, offsets computed per element (no table registers)
    const bool rowok = active & (kg >= 0) & (kg < N);
    const uint32_t base = rowok ? (uint32_t)kg * (uint32_t)(3 * n * n) + (uint32_t)rl : 0u;
    const bool first = kg == 0, last = kg == N - 1;
#pragma unroll
    for (int j = 0; j < CW; ++j) {
      const int col = cl * CW + j;
      const int seg = col / nP, cc = col - seg * nP;
      const bool ok = rowok & (seg < 3) & (cc < n) & !(first & (seg == 0)) & !(last & (seg == 2));
      const T v = gptr(A)[ok ? base + (uint32_t)((seg * n + cc) * n) : 0u];
      dst[j] = ok ? v : (T)0;
    }
  } else {
    // branch-free: every lane loads (an absent element from an in-bounds offset:
    // row rl of column 0 of its block row, or offset 0 for a row outside the
    // system) and a select zeroes what it must not hold, so the loads issue back
    // to back instead of one exec-mask branch each.  Column offsets and the masks
    // of valid / first-block / last-block columns come from compile-time tables of
    // the (at most two) chunks; 32-bit offsets (3 n^2 N < 2^31 for every size the
    // planner accepts).
    static_assert(3 * nP <= 2 * CW && CW <= 32, "branch-free loads: two chunks of at most 32 columns");
    constexpr ChunkTables<NS, CW> tab = chunk_tables<NS, CW>();
    const bool rowok = active & (kg >= 0) & (kg < N);
    // byte offsets in 32 bits: one VGPR offset per load against the uniform base
    const uint32_t base = rowok ? ((uint32_t)kg * (uint32_t)(3 * n * n) + (uint32_t)rl) * (uint32_t)sizeof(T) : 0u;
    const bool c1 = cl != 0;
    uint32_t mask = c1 ? tab.valid[1] : tab.valid[0];
    if (kg == 0) mask &= ~(c1 ? tab.seg0[1] : tab.seg0[0]);
    if (kg == N - 1) mask &= ~(c1 ? tab.seg2[1] : tab.seg2[0]);
    mask = rowok ? mask : 0u;
    const auto *bytes = reinterpret_cast<const __attribute__((address_space(1))) unsigned char *>(gptr(A));
#pragma unroll
    for (int j = 0; j < CW; ++j) {
      const uint32_t off = base + (c1 ? tab.off[1][j] : tab.off[0][j]) * (uint32_t)sizeof(T);
      const T v = *reinterpret_cast<const __attribute__((address_space(1))) T *>(bytes + off);
      dst[j] = ((mask >> j) & 1u) ? v : (T)0;
    }
  }
}

// Row-broadcast layout: the whole row `row` of block row kg (3n values, block
// column s, column k at index s*n + k).
template <typename T, int NS, bool GUARDED = false>
__device__ __forceinline__ void load_row(T (&dst)[3 * NS], const T *__restrict__ A, int kg, int row, int N,
                                         bool active) {
  constexpr int n = NS;
  int rl = row;
  asm volatile("" : "+v"(rl));
  if constexpr (GUARDED) {
#pragma unroll
    for (int j = 0; j < 3 * n; ++j) {
      const int seg = j / n;
      T v = 0;
      if (active && kg >= 0 && kg < N && !(kg == 0 && seg == 0) && !(kg == N - 1 && seg == 2))
        v = gptr(A)[(size_t)kg * 3 * n * n + (size_t)j * n + rl];
      dst[j] = v;
    }
    return;
  }
  // branch-free, as load_chunk
  const bool rowok = active & (kg >= 0) & (kg < N);
  const uint32_t base = rowok ? (uint32_t)kg * (uint32_t)(3 * n * n) + (uint32_t)rl : 0u;
  const bool first = kg == 0, last = kg == N - 1;
#pragma unroll
  for (int j = 0; j < 3 * n; ++j) {
    const int seg = j / n;
    const bool ok = rowok & !(first & (seg == 0)) & !(last & (seg == 2));
    const T v = gptr(A)[ok ? base + (uint32_t)(j * n) : 0u];
    dst[j] = ok ? v : (T)0;
  }
}

// a0/a1 += m[k] * (x of lane k of this lane's 16-lane row), k = 0..NS-1, on
// alternate accumulators.  DPP row_newbcast:k hands every lane of a row the
// value of its lane k, so one VALU op does one multiply-add with no LDS traffic.
// The compiler does not fold a broadcast into an accumulating FMA by itself,
// hence the asm; each block opens with s_nop 4, the 5 wait states a DPP read
// needs after a VALU write of EXEC (2 after a VALU write of its source VGPR).
#define GBDPCG_RC(OP, k, a, mi) OP " %" #a ", %2, %" #mi " row_newbcast:" #k " row_mask:0xf bank_mask:0xf\n\t"
#define GBDPCG_RC12(OP)                                                                                     \
  GBDPCG_RC(OP, 0, 0, 3) GBDPCG_RC(OP, 1, 1, 4) GBDPCG_RC(OP, 2, 0, 5) GBDPCG_RC(OP, 3, 1, 6)                \
  GBDPCG_RC(OP, 4, 0, 7) GBDPCG_RC(OP, 5, 1, 8) GBDPCG_RC(OP, 6, 0, 9) GBDPCG_RC(OP, 7, 1, 10)               \
  GBDPCG_RC(OP, 8, 0, 11) GBDPCG_RC(OP, 9, 1, 12) GBDPCG_RC(OP, 10, 0, 13) GBDPCG_RC(OP, 11, 1, 14)
#define GBDPCG_RC14(OP) GBDPCG_RC12(OP) GBDPCG_RC(OP, 12, 0, 15) GBDPCG_RC(OP, 13, 1, 16)
#define GBDPCG_M12(m) "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), \
    "v"(m[7]), "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11])
#define GBDPCG_M14(m) GBDPCG_M12(m), "v"(m[12]), "v"(m[13])

template <typename T, int NS>
struct RowCast;
template <>
struct RowCast<float, 14> {
  static __device__ __forceinline__ void fma(float &a0, float &a1, float x, const float *m) {
    asm("s_nop 4\n\t" GBDPCG_RC14("v_fmac_f32_dpp") : "+v"(a0), "+v"(a1) : "v"(x), GBDPCG_M14(m));
  }
};
template <>
struct RowCast<float, 12> {
  static __device__ __forceinline__ void fma(float &a0, float &a1, float x, const float *m) {
    asm("s_nop 4\n\t" GBDPCG_RC12("v_fmac_f32_dpp") : "+v"(a0), "+v"(a1) : "v"(x), GBDPCG_M12(m));
  }
};
template <>
struct RowCast<double, 14> {
  static __device__ __forceinline__ void fma(double &a0, double &a1, double x, const double *m) {
    asm("s_nop 4\n\t" GBDPCG_RC14("v_fmac_f64_dpp") : "+v"(a0), "+v"(a1) : "v"(x), GBDPCG_M14(m));
  }
};
template <>
struct RowCast<double, 12> {
  static __device__ __forceinline__ void fma(double &a0, double &a1, double x, const double *m) {
    asm("s_nop 4\n\t" GBDPCG_RC12("v_fmac_f64_dpp") : "+v"(a0), "+v"(a1) : "v"(x), GBDPCG_M12(m));
  }
};
#undef GBDPCG_RC
#undef GBDPCG_RC12
#undef GBDPCG_RC14
#undef GBDPCG_M12
#undef GBDPCG_M14

// y = row . (xa | xb | xc), x* the lane's entries of three consecutive knots;
// four accumulator chains summed in a fixed order.
template <typename T, int NS>
__device__ __forceinline__ T rowcast_dot(const T (&mat)[3 * NS], T xa, T xb, T xc) {
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  RowCast<T, NS>::fma(a0, a1, xa, mat);
  RowCast<T, NS>::fma(a2, a3, xb, mat + NS);
  RowCast<T, NS>::fma(a0, a1, xc, mat + 2 * NS);
  return (a0 + a1) + (a2 + a3);
}

__device__ __forceinline__ bool finite_d(double v) { return __builtin_isfinite(v); }
__device__ __forceinline__ bool finite_t(float v) { return __builtin_isfinite(v); }
__device__ __forceinline__ bool finite_t(double v) { return __builtin_isfinite(v); }


// One matvec row pass: this lane's row of a resident block row times the vector
// window of knots m-1..m+1 (local knots m..m+2) of `vec`; 0 on inactive lanes.
template <typename T, int NS, int NT, bool BC, int CW>
__device__ __forceinline__ T row_mv(const RegItem<T, NS, NT, BC> &it, const T (&mat)[CW], const T *vec, bool act) {
  using R = RegItem<T, NS, NT, BC>;
  if constexpr (BC) {
    // every lane of the 16-lane row takes part in the broadcasts; lanes past the
    // window read (and discard) the last knot's entries
    const int mm = it.m < it.K + 1 ? it.m : it.K + 1;
    const T *v = vec + mm * R::nP + it.row;
    const T y = rowcast_dot<T, NS>(mat, v[0], v[R::nP], v[2 * R::nP]);
    return act ? y : (T)0;
  } else {
    T y = act ? chunk_dot<T, CW>(mat, vec + it.m * R::nP + it.c * CW) : (T)0;
    return group_sum<R::LPR>(y);
  }
}

// Whether a paired instance keeps its Pinv chunks in LDS: the FP32 column-chunk
// instances of n <= 14 at 512 threads, held to 128 VGPRs so that two workgroups
// share a CU (with the chunks in registers they spill ~100 B per lane).  The
// host sizes LDS with the same rule (reg_lds_bytes).
__host__ __device__ constexpr bool pair_pinv_lds(size_t tsize, int nt, bool bc, int ns) {
  return tsize == 4 && nt == 512 && !bc && ns <= 14;
}

// Whether a paired instance keeps its S chunks in LDS: FP64 rows of n = 28 at
// 512 threads (two lanes per row, two waves per SIMD; with both matrices in
// registers a lane would need ~290 registers).  Slots for the owned rows only:
// pair_s_slots(nt, n) lanes (two ghost knots' lanes fewer than the workgroup).
// Such an instance runs paired items only (plans of forward + adjoint systems).
// The host sizes LDS with the same rules (reg_lds_bytes).
__host__ __device__ constexpr bool pair_s_lds(size_t tsize, int nt, bool bc, int ns) {
  return tsize == 8 && nt == 512 && !bc && ns == 28;
}
__host__ __device__ constexpr int pair_s_slots(int nt, int ns) { return nt - 2 * 2 * ns; }

// u = Pinv r of both sides of a paired solve: from the register chunk, or from
// the LDS chunk (PLDS).
template <bool PLDS, typename T, int NS, int NT, bool BC, int CWR>
__device__ __forceinline__ void pinv_mv2(const RegItem<T, NS, NT, BC> &it, const T (&mM)[CWR], const T *plds,
                                         const T *va, const T *vb, int tid, T &ya, T &yb) {
  using R = RegItem<T, NS, NT, BC>;
  if constexpr (PLDS) {
    static_assert(sizeof(T) == 4 && !BC, "LDS-resident Pinv: FP32 column chunks only");
    T a = 0, b = 0;
    if (it.act_u) {
      const int off = it.m * R::nP + it.c * R::CW;
      chunk_dot2_lds<R::CW, NT>(plds, tid, va + off, vb + off, a, b);
    }
    ya = group_sum<R::LPR>(a);
    yb = group_sum<R::LPR>(b);
  } else {
    ya = row_mv(it, mM, va, it.act_u);
    yb = row_mv(it, mM, vb, it.act_u);
  }
}

// S times one or two vectors in reg_item_pair: from the register chunk, or
// from the LDS chunk (SLDS: slot = thread index less the left ghost knot's
// lanes).
template <bool SLDS, int SL, typename T, int NS, int NT, bool BC, int CWR>
__device__ __forceinline__ T s_mv1(const RegItem<T, NS, NT, BC> &it, const T (&mS)[CWR], const T *slds, const T *v,
                                   int tid) {
  using R = RegItem<T, NS, NT, BC>;
  if constexpr (SLDS) {
    T a = 0;
    if (it.act_w) a = chunk_dot_lds_d<R::CW, SL>(slds, tid - R::LPR * NS, v + it.m * R::nP + it.c * R::CW);
    return group_sum<R::LPR>(a);
  } else {
    return row_mv(it, mS, v, it.act_w);
  }
}

template <bool SLDS, int SL, typename T, int NS, int NT, bool BC, int CWR>
__device__ __forceinline__ void s_mv2(const RegItem<T, NS, NT, BC> &it, const T (&mS)[CWR], const T *slds,
                                      const T *va, const T *vb, int tid, T &ya, T &yb) {
  using R = RegItem<T, NS, NT, BC>;
  if constexpr (SLDS) {
    T a = 0, b = 0;
    if (it.act_w) {
      const int off = it.m * R::nP + it.c * R::CW;
      chunk_dot2_lds_d<R::CW, SL>(slds, tid - R::LPR * NS, va + off, vb + off, a, b);
    }
    ya = group_sum<R::LPR>(a);
    yb = group_sum<R::LPR>(b);
  } else {
    ya = row_mv(it, mS, va, it.act_w);
    yb = row_mv(it, mS, vb, it.act_w);
  }
}

// g_S block rows of a rank's owned knots a..a+K-1 (paired items): -1/2 (gg_k lam_j^T +
// lam_k gg_j^T) for j in {k-1, k, k+1}, zero outside the system.  gg and lam are
// the ghost-extended windows in LDS (knot a-1+i at (i+1) nP).  V consecutive rows
// of one block column are contiguous in the cbtd layout, so each thread forms V
// of them and stores them at once (16 bytes where n and gS's alignment allow).
template <typename T, int NS, int NT>
__device__ __forceinline__ void write_gS(T *gS, const T *gg, const T *lam, int a, int K, int N) {
  constexpr int n = NS, nP = (int)padded_n(NS), bs = 3 * NS * NS;
  constexpr int VM = (int)(16 / sizeof(T));
  constexpr int V = n % VM == 0 ? VM : (n % 2 == 0 ? 2 : 1);
  typedef T vec_t __attribute__((ext_vector_type(V)));
  const bool vec_ok = V > 1 && ((uintptr_t)gS & (V * sizeof(T) - 1)) == 0;
  const int step = vec_ok ? V : 1;
  for (int e = opaque_tid() * step; e < K * bs; e += NT * step) {
    const int kl = e / bs;
    const int wq = e - kl * bs;
    const int col = wq / n;
    const int rr_ = wq - col * n;
    const int seg = col / n;
    const int cc = col - seg * n;
    const int k = a + kl;
    const int j = k - 1 + seg;
    const bool in = j >= 0 && j < N;
    const int lk = (kl + 2) * nP, lj = (kl + 1 + seg) * nP;
    const T lj_v = in ? lam[lj + cc] : (T)0, gj_v = in ? gg[lj + cc] : (T)0;
    if (vec_ok) {
      vec_t out;
#pragma unroll
      for (int v = 0; v < V; ++v) out[v] = in ? (T)-0.5 * (gg[lk + rr_ + v] * lj_v + lam[lk + rr_ + v] * gj_v) : (T)0;
      *reinterpret_cast<__attribute__((address_space(1))) vec_t *>(gptr(gS) + (size_t)a * bs + e) = out;
    } else {
      gptr(gS)[(size_t)a * bs + e] = in ? (T)-0.5 * (gg[lk + rr_] * lj_v + lam[lk + rr_] * gj_v) : (T)0;
    }
  }
}

// PRE: the set-up's HBM reads (initial guess window, first right-hand side)
// issued together with the matrix chunks, one round trip instead of three (off
// for the one-pass items of the paired kernel, whose 128-VGPR allocation spills
// differently, and runs its batches slower, with it; the paired item issues its
// own set-up reads together).
// BF: the matrix chunks loaded branch-free (load_chunk), for the fixed-size
// instances; the runtime-n instances keep the guarded loads.
template <typename T, int NS, int NT, bool FUSED, bool BC = false, bool PRE = true, bool BF = PRE>
__device__ __forceinline__ void reg_item(const KSys &d, uint32_t rank, T *lds) {
  using R = RegItem<T, NS, NT, BC>;
  constexpr int n = R::n, nP = R::nP, CW = R::CW;
  R it;
  it.N = (int)d.N;
  it.G = (int)d.G;
  it.rank = (int)rank;
  it.a = (int)knot_begin(rank, d.G, d.N);
  it.b = (int)knot_begin(rank + 1, d.G, d.N);
  it.K = it.b - it.a;
  const int K = it.K;
  it.VL = (K + 4) * nP + (BC ? 16 : R::LPR * CW);
  it.slots = d.slots;
  it.tol = d.tol;
  it.rel = (d.flags & KF_REL) != 0;
  it.xcd = false;
  it.xcd_ok = (d.flags & KF_NOXCD) == 0;
  // lane roles (from an opaque thread index: recomputed per item rather than
  // hoisted out of the batched item loop)
  const int t = opaque_tid();
  it.q = t / R::LPR;
  it.c = t % R::LPR;
  if constexpr (BC) {
    it.m = t >> 4;
    it.row = t & 15;
  } else {
    it.m = it.q / n;
    it.row = it.q - it.m * n;
  }
  it.kg = it.a - 1 + it.m;
  it.act_u = it.m < K + 2 && (!BC || it.row < n);
  it.act_w = it.m >= 1 && it.m <= K && (!BC || it.row < n);
  it.owner = it.act_w && it.c == 0;
  it.ghost = it.c == R::LPR - 1 && it.q < 4 * n;
  it.set_edge_slots();
  {
    const int gk = it.q / n, gr = it.q - (it.q / n) * n;  // ghost element q: knot slot 0..3
    const int gj = gk < 2 ? gk : K + gk;                  // local knots 0, 1, K+2, K+3
    it.gpos = gj * nP + gr;
  }
  // LDS carve-up
  it.rv = lds;
  it.uv = it.rv + it.VL;
  it.wv = it.uv + it.VL;
  it.xv = it.wv + it.VL;
  it.lv = it.xv + it.VL;
  it.wsum = it.lv + it.VL;
  it.tot = it.wsum + 3 * R::NWAVES;
  it.ctl = reinterpret_cast<int *>(it.tot + 4);
  const int total = (int)(it.tot + 4 - lds);

  const int m = it.m, row = it.row, c = it.c, kg = it.kg;
  const int own_pos = (m + 1) * nP + row;  // local knot of own row kl = m - 1 is m + 1

  if (t == 0) {
    const uint32_t ea = (uint32_t)(xload(it.rec(0, it.rank)) >> 32);
    const uint32_t eb = (uint32_t)(xload(it.rec(1, it.rank)) >> 32);
    uint32_t base = (ea > eb ? ea : eb) + 1u;
    if (base > 0xFFF00000u) base = 1u;
    it.ctl[0] = 0;
    it.ctl[1] = (int)base;
  }
  for (int i = opaque_tid(); i < total; i += NT) lds[i] = (T)0;

  const bool fwd = (d.flags & KF_FWD) != 0;
  const bool adj = (d.flags & KF_ADJ) != 0;
  // every HBM read of the set-up in one round trip (PRE): the initial guess
  // window (own + one ghost knot each side; (K + 2) n <= NT values, one per
  // thread below NT, the rest after the barrier) and the first right-hand side,
  // issued with the matrix chunks
  const T *lam0 = static_cast<const T *>(d.lambda);
  const int wl = (K + 2) * n;
  T lam_pre = 0;
  int lam_pos = -1;
  if constexpr (PRE) {
    const int i = opaque_tid();
    if (i < wl) {
      const int jj = i / n, e = i - jj * n;
      const int k = it.a - 1 + jj;
      if (k >= 0 && k < it.N) {
        lam_pre = gptr(lam0)[(size_t)k * n + e];
        lam_pos = (jj + 1) * nP + e;
      }
    }
  }
  T b_pre = 0;
  if constexpr (PRE) {
    if (it.owner) b_pre = gptr(static_cast<const T *>(fwd ? d.gamma : d.glambda))[(size_t)kg * n + row];
  }

  // matrix chunks straight from HBM into registers
  T mM[CW], mS[CW];
  // branch-free loads for the fixed-size instances of n <= 14 (the single-solve
  // layouts); the runtime-n and n = 28 instances keep the guarded form, whose
  // fewer live masks leave their batches faster
  constexpr bool kGuarded = !(BF && NS <= 14);
  constexpr int kMode = kGuarded ? (R::LPR == 1 && CW == 3 * nP ? kLoadRows : kLoadGuarded)
                                 : (sizeof(T) == 4 && NT == 256) ? kLoadTables : kLoadArith;
  // S first: loads complete in issue order, so the first product r = b - S x
  // (and the first exchange) can start while the Pinv chunks are in flight
  if constexpr (BC) {
    load_row<T, NS, kGuarded>(mS, static_cast<const T *>(d.S), kg, row, it.N, it.act_w);
    load_row<T, NS, kGuarded>(mM, static_cast<const T *>(d.Pinv), kg, row, it.N, it.act_u);
  } else {
    load_chunk<T, NS, CW, kMode>(mS, static_cast<const T *>(d.S), kg, row, c, it.N, it.act_w);
    load_chunk<T, NS, CW, kMode>(mM, static_cast<const T *>(d.Pinv), kg, row, c, it.N, it.act_u);
  }
  __syncthreads();
  it.base = (uint32_t)it.ctl[1];
  it.seq = 0;
  {
    T *dst = fwd ? it.xv : it.lv;
    if (lam_pos >= 0) dst[lam_pos] = lam_pre;
    for (int i = opaque_tid() + (PRE ? NT : 0); i < wl; i += NT) {
      const int jj = i / n, e = i - jj * n;
      const int k = it.a - 1 + jj;
      if (k >= 0 && k < it.N) dst[(jj + 1) * nP + e] = gptr(lam0)[(size_t)k * n + e];
    }
  }
  __syncthreads();

  RegOut fo{0, GBDPCG_CONVERGED, 0.0, 0.0}, ao{0, GBDPCG_CONVERGED, 0.0, 0.0};
  bool timed_out = false;
  for (int pass = fwd ? 0 : 1; pass <= (adj ? 1 : 0); ++pass) {
    RegOut out{0, GBDPCG_MAX_ITER, 0.0, 0.0};
    if (pass == 1) {
      // adjoint: lambda (own + ghosts) was left in xv by the forward pass, or
      // loaded into lv above; keep it in lv and restart from x = 0
      if (fwd) {
        for (int i = opaque_tid(); i < it.VL; i += NT) it.lv[i] = it.xv[i];
      }
      for (int i = opaque_tid(); i < it.VL; i += NT) it.xv[i] = (T)0;
      __syncthreads();
    }
    // with PRE the first pass's right-hand side was read with the set-up
    const T *rhs = static_cast<const T *>(pass == 0 ? d.gamma : d.glambda);
    T x_i = 0, b_i = 0;
    if (it.owner) {
      x_i = it.xv[own_pos];
      b_i = PRE && (pass == 0) == fwd ? b_pre : gptr(rhs)[(size_t)kg * n + row];
    }
    // r = rhs - S x (own rows)
    T r_i;
    {
      T y = row_mv(it, mS, it.xv, it.act_w);
      r_i = b_i - y;
    }
    if (it.owner) it.rv[own_pos] = r_i;
    it.publish_edge(r_i, 2);
    it.reduce_publish(it.owner ? b_i * b_i : (T)0, (T)0, (T)0, true);
    T gg, sx, sxx;
    if (!it.gather(it.rv, 2, gg, sx, sxx)) {
      out.status = GBDPCG_TIMED_OUT;
      timed_out = true;
      if (pass == 0) fo = out; else ao = out;
      break;
    }
    // the sums are of small integers, exact in T; compare in double
    // (uniform by construction; readfirstlane lets the compiler branch on it in
    // scalar code)
    it.xcd = __builtin_amdgcn_readfirstlane(
                 (int)(it.xcd_ok && it.G > 1 && (double)sx * (double)sx == (double)it.G * (double)sxx)) != 0;
    // exit threshold: |r'z| < tol (reference) or r'r <= tol^2 rhs'rhs
    const double thr = it.rel ? it.tol * it.tol * (double)gg : it.tol;
    T r_g = it.ghost ? it.rv[it.gpos] : (T)0;
    if constexpr (!FUSED) {
      // reference recurrence.  z = Pinv r on own and one ghost knot each side,
      // p = z, eta = r'z
      T z_i, p_i, eta, rr;
      {
        z_i = row_mv(it, mM, it.rv, it.act_u);
        p_i = z_i;
        if (it.act_u && c == 0) it.uv[(m + 1) * nP + row] = p_i;  // uv holds p
        const T rz = it.owner ? r_i * z_i : (T)0;
        const T r2 = it.owner ? r_i * r_i : (T)0;
        it.reduce_publish(rz, (T)0, r2);
        T unused;
        if (!it.gather(it.wv, 0, eta, unused, rr)) {
          out.status = GBDPCG_TIMED_OUT;
          timed_out = true;
          if (pass == 0) fo = out; else ao = out;
          break;
        }
        out.eta = (double)eta;
        out.rr = (double)rr;
        bool stop = false;
        if (!finite_t(eta) || !finite_t(rr)) {
          out.status = GBDPCG_NONFINITE;
          stop = true;
        } else if (eta == (T)0) {
          out.status = rr == (T)0 ? GBDPCG_CONVERGED : GBDPCG_BREAKDOWN;
          stop = true;
        }
        if (stop) {
          out.iters = 0;
          goto finish;
        }
      }
      {
        uint32_t iter = 0;
        const uint32_t max_iter = d.max_iter;
        while (iter < max_iter) {
          // v = S p (own rows); alpha = eta / p'v.  The exchange also carries the
          // two edge knots of v each side: the ghost lanes' r needs them.
          T v_i;
          {
            v_i = row_mv(it, mS, it.uv, it.act_w);
          }
          it.publish_edge(v_i, 2);
          it.reduce_publish(it.owner ? p_i * v_i : (T)0, (T)0, (T)0);
          T pv, u1, u2;
          if (!it.gather(it.wv, 2, pv, u1, u2)) {
            out.status = GBDPCG_TIMED_OUT;
            timed_out = true;
            break;
          }
          const T v_g = it.ghost ? it.wv[it.gpos] : (T)0;
          if (pv == (T)0 || !finite_t(pv)) {
            out.status = pv == (T)0 ? GBDPCG_BREAKDOWN : GBDPCG_NONFINITE;
            break;
          }
          const T alpha = qdiv(eta, pv);
          // lambda += alpha p ; r -= alpha v (own rows, and the ghost lanes' r)
          if (it.owner) {
            x_i = fmaT(alpha, p_i, x_i);
            r_i = fmaT(-alpha, v_i, r_i);
            it.rv[own_pos] = r_i;
          }
          if (it.ghost) {
            r_g = fmaT(-alpha, v_g, r_g);
            it.rv[it.gpos] = r_g;
          }
          __syncthreads();
          // z = Pinv r (own + ghost rows) ; eta' = r'z ; exit test
          {
            z_i = row_mv(it, mM, it.rv, it.act_u);
          }
          const T rz = it.owner ? r_i * z_i : (T)0;
          const T r2 = it.owner ? r_i * r_i : (T)0;
          it.reduce_publish(rz, (T)0, r2);
          T eta_new, u3;
          if (!it.gather(it.wv, 0, eta_new, u3, rr)) {
            out.status = GBDPCG_TIMED_OUT;
            timed_out = true;
            break;
          }
          ++iter;
          out.eta = (double)eta_new;
          out.rr = (double)rr;
          if (it.rel ? ((double)rr <= thr) : (__builtin_fabs((double)eta_new) < thr)) {
            out.status = GBDPCG_CONVERGED;
            break;
          }
          if (!finite_t(eta_new)) {
            out.status = GBDPCG_NONFINITE;
            break;
          }
          // p = z + beta p on own and ghost knots
          const T beta = qdiv(eta_new, eta);
          eta = eta_new;
          p_i = fmaT(beta, p_i, z_i);
          if (it.act_u && c == 0) it.uv[(m + 1) * nP + row] = p_i;
          __syncthreads();
        }
        out.iters = (int)iter;
      }
    } else {
      // u = Pinv r (own + ghost rows), w = S u (own rows)
      T u_i, w_i;
      T p_i, s_i, s_g;
      T eta, alpha, rr;
      {
        u_i = row_mv(it, mM, it.rv, it.act_u);
        if (it.act_u && c == 0) it.uv[(m + 1) * nP + row] = u_i;
        __syncthreads();
        w_i = row_mv(it, mS, it.uv, it.act_w);
        it.publish_edge(w_i, 2);
        const T rz = it.owner ? r_i * u_i : (T)0;
        const T wu = it.owner ? w_i * u_i : (T)0;
        const T r2 = it.owner ? r_i * r_i : (T)0;
        it.reduce_publish(rz, wu, r2);
        T delta;
        if (!it.gather(it.wv, 2, eta, delta, rr)) {
          out.status = GBDPCG_TIMED_OUT;
          timed_out = true;
          if (pass == 0) fo = out; else ao = out;
          break;
        }
        p_i = u_i;
        s_i = w_i;
        s_g = it.ghost ? it.wv[it.gpos] : (T)0;
        alpha = delta != (T)0 ? eta / delta : (T)0;
        out.eta = (double)eta;
        out.rr = (double)rr;
        bool stop = false;
        if (!finite_d((double)eta) || !finite_d((double)rr)) {
          out.status = GBDPCG_NONFINITE;
          stop = true;
        } else if (eta == (T)0) {
          out.status = rr == (T)0 ? GBDPCG_CONVERGED : GBDPCG_BREAKDOWN;
          stop = true;
        } else if (delta == (T)0 || !finite_d((double)alpha)) {
          out.status = GBDPCG_BREAKDOWN;
          stop = true;
        }
        if (stop) {
          out.iters = 0;
          goto finish;
        }
      }
      {
        uint32_t iter = 0;
        const uint32_t max_iter = d.max_iter;
        T last_eta = eta, last_rr = rr;  // of the last exchange (uniform)
        while (iter < max_iter) {
          // x += alpha p ; r -= alpha s  (own in registers, ghosts by the ghost lanes)
          if (it.owner) {
            x_i = fmaT(alpha, p_i, x_i);
            r_i = fmaT(-alpha, s_i, r_i);
            it.rv[own_pos] = r_i;
          }
          if (it.ghost) {
            r_g = fmaT(-alpha, s_g, r_g);
            it.rv[it.gpos] = r_g;
          }
          __syncthreads();
          u_i = row_mv(it, mM, it.rv, it.act_u);
          if (it.act_u && c == 0) it.uv[(m + 1) * nP + row] = u_i;
          __syncthreads();
          w_i = row_mv(it, mS, it.uv, it.act_w);
          it.publish_edge(w_i, 2);
          const T rz = it.owner ? r_i * u_i : (T)0;
          const T wu = it.owner ? w_i * u_i : (T)0;
          const T r2 = it.owner ? r_i * r_i : (T)0;
          it.reduce_publish(rz, wu, r2);
          T eta_new, delta_new;
          if (!it.gather(it.wv, 2, eta_new, delta_new, rr)) {
            out.status = GBDPCG_TIMED_OUT;
            timed_out = true;
            break;
          }
          // ghost lanes: the neighbour's w, read now so the LDS latency overlaps the
          // scalar recurrence below
          const T w_g = it.ghost ? it.wv[it.gpos] : (T)0;
          ++iter;
          last_eta = eta_new;
          last_rr = rr;
          // the stop tests and the next coefficients are independent: all are
          // computed first, then one branch (a shorter dependency chain; the
          // coefficients of a stopping iteration are discarded).  beta and
          // eta'/alpha: two divisions in flight at once.
          const bool conv = it.rel ? ((double)rr <= thr) : (__builtin_fabs((double)eta_new) < thr);
          const bool nonfin = !finite_t(eta_new);
          const T beta = qdiv(eta_new, eta);
          const T den = delta_new - beta * qdiv(eta_new, alpha);
          const bool brk = den == (T)0 || !finite_t(den);
          const T alpha_new = qdiv(eta_new, den);
          if (conv || nonfin || brk) {
            out.status = conv ? GBDPCG_CONVERGED : nonfin ? GBDPCG_NONFINITE : GBDPCG_BREAKDOWN;
            break;
          }
          alpha = alpha_new;
          eta = eta_new;
          p_i = fmaT(beta, p_i, u_i);
          s_i = fmaT(beta, s_i, w_i);
          if (it.ghost) s_g = fmaT(beta, s_g, w_g);
        }
        out.iters = (int)iter;
        out.eta = (double)last_eta;
        out.rr = (double)last_rr;
      }
    }
  finish:
    if (pass == 0) fo = out; else ao = out;
    if (timed_out) break;
    // solution: own knots to HBM; for the adjoint, also its edge knots to the
    // neighbours (g_S needs lambda and g_gamma one knot beyond the owned range)
    {
      T *dst = static_cast<T *>(pass == 0 ? d.lambda : d.ggamma);
      if (it.owner) gptr(dst)[(size_t)kg * n + row] = x_i;
    }
    if (adj) {
      if (it.owner) it.xv[own_pos] = x_i;
      it.publish_edge(x_i, 1);
      it.reduce_publish((T)0, (T)0, (T)0);
      T z0, z1, z2;
      if (!it.gather(it.xv, 1, z0, z1, z2)) {
        timed_out = true;
        if (pass == 0) fo.status = GBDPCG_TIMED_OUT; else ao.status = GBDPCG_TIMED_OUT;
        break;
      }
    }
    if (pass == 1) {
      // g_S block rows of the own knots: -1/2 (gg_k lam_j^T + lam_k gg_j^T), j in {k-1,k,k+1}
      // (one element per thread and step: the vectorised write_gS costs this
      // item two VGPRs, and with them the 512-thread FP32 n = 14 instance's
      // second workgroup per CU)
      const int bs = 3 * n * n;
      T *gS = static_cast<T *>(d.gS);
      for (int e = opaque_tid(); e < K * bs; e += NT) {
        const int kl = e / bs;
        const int wq = e - kl * bs;
        const int col = wq / n;
        const int rr_ = wq - col * n;
        const int seg = col / n;
        const int cc = col - seg * n;
        const int k = it.a + kl;
        const int j = k - 1 + seg;
        T val = 0;
        if (j >= 0 && j < it.N) {
          const int lk = (kl + 2) * nP, lj = (kl + 1 + seg) * nP;
          val = (T)-0.5 * (it.xv[lk + rr_] * it.lv[lj + cc] + it.lv[lk + rr_] * it.xv[lj + cc]);
        }
        gptr(gS)[(size_t)it.a * bs + e] = val;
      }
    }
    __syncthreads();
  }
  if (rank == 0 && t == 0 && d.result != nullptr) {
    gbdpcg_result res;
    res.iters = fo.iters;
    res.status = fo.status;
    res.adj_iters = ao.iters;
    res.adj_status = ao.status;
    res.eta = fo.eta;
    res.rr = fo.rr;
    res.adj_eta = ao.eta;
    res.adj_rr = ao.rr;
    *d.result = res;
  }
  if (timed_out && t == 0 && d.err != nullptr)
    __hip_atomic_store(d.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Forward and adjoint solves of one system iterated together (fused recurrence;
// both flags set).  The two solves share S and Pinv and are independent until
// g_S is formed, so each iteration runs the two right-hand sides (A: gamma,
// from lambda0; B: g_lambda, from 0) through the same resident matrices and one
// exchange carrying both sets of scalars and edges.  Each side performs exactly
// the operations of its own solve in reg_item, in the same order, and stops on
// its own exit test; once one side has stopped, the other runs alone (its
// matvecs and updates are skipped, the exchange continues for the other).
template <typename T, int NS, int NT, bool BC>
__device__ __forceinline__ void reg_item_pair(const KSys &d, uint32_t rank, T *lds) {
  using R = RegItem<T, NS, NT, BC>;
  constexpr int n = R::n, nP = R::nP, CW = R::CW;
  // Pinv chunks in LDS (the first NT * CW values) for the instances held to
  // 128 VGPRs; the vectors follow
  constexpr bool PLDS = pair_pinv_lds(sizeof(T), NT, BC, NS);
  T *plds = nullptr;
  if constexpr (PLDS) {
    plds = lds;
    lds += (size_t)NT * CW;
  }
  // S chunks in LDS (the first pair_s_slots * CW values) for the FP64 n = 28
  // instance at 512 threads; its vectors need only 8 values of read-ahead
  constexpr bool SLDS = pair_s_lds(sizeof(T), NT, BC, NS);
  constexpr int SL = SLDS ? pair_s_slots(NT, NS) : 1;
  T *slds = nullptr;
  if constexpr (SLDS) {
    slds = lds;
    lds += (size_t)SL * CW;
  }
  R it;
  it.N = (int)d.N;
  it.G = (int)d.G;
  it.rank = (int)rank;
  it.a = (int)knot_begin(rank, d.G, d.N);
  it.b = (int)knot_begin(rank + 1, d.G, d.N);
  it.K = it.b - it.a;
  const int K = it.K;
  it.VL = (K + 4) * nP + (BC ? 16 : SLDS ? 8 : R::LPR * CW);
  it.slots = d.slots;
  it.tol = d.tol;
  it.rel = (d.flags & KF_REL) != 0;
  it.xcd = false;
  it.xcd_ok = (d.flags & KF_NOXCD) == 0;
  const int t = opaque_tid();
  it.q = t / R::LPR;
  it.c = t % R::LPR;
  if constexpr (BC) {
    it.m = t >> 4;
    it.row = t & 15;
  } else {
    it.m = it.q / n;
    it.row = it.q - it.m * n;
  }
  it.kg = it.a - 1 + it.m;
  it.act_u = it.m < K + 2 && (!BC || it.row < n);
  it.act_w = it.m >= 1 && it.m <= K && (!BC || it.row < n);
  it.owner = it.act_w && it.c == 0;
  it.ghost = it.c == R::LPR - 1 && it.q < 4 * n;
  it.set_edge_slots();
  {
    const int gk = it.q / n, gr = it.q - (it.q / n) * n;
    const int gj = gk < 2 ? gk : K + gk;
    it.gpos = gj * nP + gr;
  }
  // LDS: r, u, w of each side, lambda and g_gamma windows, 6 partials per wave
  T *rA = lds, *uA = rA + it.VL, *wA = uA + it.VL;
  T *rB = wA + it.VL, *uB = rB + it.VL, *wB = uB + it.VL;
  it.xv = wB + it.VL;  // lambda0 window, later g_gamma with ghosts
  it.lv = it.xv + it.VL;  // lambda with ghosts (for g_S)
  it.rv = rA;
  it.uv = uA;
  it.wv = wA;
  it.wsum = it.lv + it.VL;
  it.tot = it.wsum + 6 * R::NWAVES;
  it.ctl = reinterpret_cast<int *>(it.tot + 8);
  const int total = (int)(it.tot + 8 - lds);

  const int m = it.m, row = it.row, c = it.c, kg = it.kg;
  const int own_pos = (m + 1) * nP + row;
  constexpr int VB = 4 * n;  // second edge block

  if (t == 0) {
    const uint32_t ea = (uint32_t)(xload(it.rec(0, it.rank)) >> 32);
    const uint32_t eb = (uint32_t)(xload(it.rec(1, it.rank)) >> 32);
    uint32_t base = (ea > eb ? ea : eb) + 1u;
    if (base > 0xFFF00000u) base = 1u;
    it.ctl[0] = 0;
    it.ctl[1] = (int)base;
  }
  for (int i = opaque_tid(); i < total; i += NT) lds[i] = (T)0;

  // the set-up's vector reads (initial guess window: (K + 2) n <= NT values, one
  // per thread; both right-hand sides of the owned rows) issued with the matrix
  // chunks: one HBM round trip instead of three
  const T *lam0 = static_cast<const T *>(d.lambda);
  const int wl = (K + 2) * n;
  T lam_pre = 0, bA = 0, bB = 0;
  int lam_pos = -1;
  {
    const int i = opaque_tid();
    if (i < wl) {
      const int jj = i / n, e = i - jj * n;
      const int k = it.a - 1 + jj;
      if (k >= 0 && k < it.N) {
        lam_pre = gptr(lam0)[(size_t)k * n + e];
        lam_pos = (jj + 1) * nP + e;
      }
    }
    if (it.owner) {
      bA = gptr(static_cast<const T *>(d.gamma))[(size_t)kg * n + row];
      bB = gptr(static_cast<const T *>(d.glambda))[(size_t)kg * n + row];
    }
  }

  T mM[PLDS ? 1 : CW], mS[SLDS ? 1 : CW];
  if constexpr (BC) {
    load_row<T, NS, true>(mM, static_cast<const T *>(d.Pinv), kg, row, it.N, it.act_u);
    load_row<T, NS, true>(mS, static_cast<const T *>(d.S), kg, row, it.N, it.act_w);
  } else if constexpr (SLDS) {
    // this lane's S chunk through registers into its LDS slot (read back by this
    // lane only), then its Pinv chunk into registers
    {
      T tmp[CW];
      load_chunk<T, NS, CW, kLoadGuarded>(tmp, static_cast<const T *>(d.S), kg, row, c, it.N, it.act_w);
      if (it.act_w) {
        const int si = t - R::LPR * n;
#pragma unroll
        for (int j = 0; j < CW; j += 2)
          *reinterpret_cast<double2 *>(slds + ((size_t)(j / 2) * SL + si) * 2) = make_double2(tmp[j], tmp[j + 1]);
      }
    }
    load_chunk<T, NS, CW, kLoadGuarded>(mM, static_cast<const T *>(d.Pinv), kg, row, c, it.N, it.act_u);
  } else if constexpr (!PLDS) {
    constexpr int kPairLoad = R::LPR == 1 && CW == 3 * nP ? kLoadRows : kLoadGuarded;
    load_chunk<T, NS, CW, kPairLoad>(mM, static_cast<const T *>(d.Pinv), kg, row, c, it.N, it.act_u);
    load_chunk<T, NS, CW, kPairLoad>(mS, static_cast<const T *>(d.S), kg, row, c, it.N, it.act_w);
  } else {
    // this lane's Pinv chunk through registers into its LDS slots (read back by
    // this lane only: no barrier needed between the store and the reads)
    T tmp[CW];
    load_chunk<T, NS, CW, kLoadGuarded>(tmp, static_cast<const T *>(d.Pinv), kg, row, c, it.N, it.act_u);
#pragma unroll
    for (int j = 0; j < CW; j += 4)
      *reinterpret_cast<float4 *>(plds + ((size_t)(j / 4) * NT + t) * 4) = make_float4(tmp[j], tmp[j + 1], tmp[j + 2],
                                                                                        tmp[j + 3]);
    load_chunk<T, NS, CW, kLoadGuarded>(mS, static_cast<const T *>(d.S), kg, row, c, it.N, it.act_w);
  }
  __syncthreads();
  it.base = (uint32_t)it.ctl[1];
  it.seq = 0;
  if (lam_pos >= 0) it.xv[lam_pos] = lam_pre;
  for (int i = opaque_tid() + NT; i < wl; i += NT) {
    const int jj = i / n, e = i - jj * n;
    const int k = it.a - 1 + jj;
    if (k >= 0 && k < it.N) it.xv[(jj + 1) * nP + e] = gptr(lam0)[(size_t)k * n + e];
  }
  __syncthreads();

  RegOut fo{0, GBDPCG_MAX_ITER, 0.0, 0.0}, ao{0, GBDPCG_MAX_ITER, 0.0, 0.0};
  bool timed_out = false;
  T xA = 0, xB = 0;
  if (it.owner) xA = it.xv[own_pos];
  // r = rhs - S x (B starts from x = 0: r = g_lambda)
  T rA_i;
  if constexpr (SLDS)
    rA_i = bA - s_mv1<SLDS, SL>(it, mS, slds, it.xv, t);
  else
    rA_i = bA - row_mv(it, mS, it.xv, it.act_w);
  T rB_i = bB;
  if (it.owner) {
    rA[own_pos] = rA_i;
    rB[own_pos] = rB_i;
  }
  it.publish_edge(rA_i, 2);
  it.publish_edge(rB_i, 2, VB);
  T tt[6];
  it.reduce_publish6(it.owner ? bA * bA : (T)0, (T)0, (T)0, it.owner ? bB * bB : (T)0, (T)0, (T)0, true);
  bool runA = true, runB = true;
  double thrA = 0.0, thrB = 0.0;
  T rgA = 0, rgB = 0;
  T pA = 0, sA = 0, sgA = 0, etaA = 0, alphaA = 0;
  T pB = 0, sB = 0, sgB = 0, etaB = 0, alphaB = 0;
  T uA_i = 0, wA_i = 0, uB_i = 0, wB_i = 0;
  uint32_t iterA = 0, iterB = 0;
  const uint32_t max_iter = d.max_iter;
  if (!it.gather_pair(rA, rB, 2, tt)) {
    timed_out = true;
  } else {
    it.xcd = __builtin_amdgcn_readfirstlane(
                 (int)(it.xcd_ok && it.G > 1 && (double)tt[1] * (double)tt[1] == (double)it.G * (double)tt[2])) != 0;
    thrA = uniform(it.rel ? it.tol * it.tol * (double)tt[0] : it.tol);
    thrB = uniform(it.rel ? it.tol * it.tol * (double)tt[3] : it.tol);
    rgA = it.ghost ? rA[it.gpos] : (T)0;
    rgB = it.ghost ? rB[it.gpos] : (T)0;
    // first u = Pinv r, w = S u of both sides
    if constexpr (PLDS) {
      pinv_mv2<PLDS>(it, mM, plds, rA, rB, t, uA_i, uB_i);
    } else {
      uA_i = row_mv(it, mM, rA, it.act_u);
      uB_i = row_mv(it, mM, rB, it.act_u);
    }
    if (it.act_u && c == 0) {
      uA[(m + 1) * nP + row] = uA_i;
      uB[(m + 1) * nP + row] = uB_i;
    }
    __syncthreads();
    if constexpr (SLDS) {
      s_mv2<SLDS, SL>(it, mS, slds, uA, uB, t, wA_i, wB_i);
    } else {
      wA_i = row_mv(it, mS, uA, it.act_w);
      wB_i = row_mv(it, mS, uB, it.act_w);
    }
    it.publish_edge(wA_i, 2);
    it.publish_edge(wB_i, 2, VB);
    it.reduce_publish6(it.owner ? rA_i * uA_i : (T)0, it.owner ? wA_i * uA_i : (T)0, it.owner ? rA_i * rA_i : (T)0,
                       it.owner ? rB_i * uB_i : (T)0, it.owner ? wB_i * uB_i : (T)0, it.owner ? rB_i * rB_i : (T)0);
    if (!it.gather_pair(wA, wB, 2, tt)) timed_out = true;
  }
  // initial checks of each side (as in reg_item)
  // last r'z and r'r of each side (uniform, kept as T; widened at the end)
  T lastA[2] = {0, 0}, lastB[2] = {0, 0};
  auto first = [&](RegOut &out, bool &run, T eta, T delta, T rr, T w_i, T u_i, T *wv, T &p, T &s_, T &sg, T &et,
                   T &alpha, T (&last)[2]) {
    p = u_i;
    s_ = w_i;
    sg = it.ghost ? wv[it.gpos] : (T)0;
    et = eta;
    alpha = uniform(delta != (T)0 ? eta / delta : (T)0);
    last[0] = eta;
    last[1] = rr;
    bool stop = false;
    if (!finite_d((double)eta) || !finite_d((double)rr)) {
      out.status = GBDPCG_NONFINITE;
      stop = true;
    } else if (eta == (T)0) {
      out.status = rr == (T)0 ? GBDPCG_CONVERGED : GBDPCG_BREAKDOWN;
      stop = true;
    } else if (delta == (T)0 || !finite_d((double)alpha)) {
      out.status = GBDPCG_BREAKDOWN;
      stop = true;
    }
    if (stop) run = false;
    if (max_iter == 0) run = false;
  };
  if (!timed_out) {
    first(fo, runA, tt[0], tt[1], tt[2], wA_i, uA_i, wA, pA, sA, sgA, etaA, alphaA, lastA);
    first(ao, runB, tt[3], tt[4], tt[5], wB_i, uB_i, wB, pB, sB, sgB, etaB, alphaB, lastB);
  } else {
    runA = runB = false;
  }
  // The scalar step of one side after an exchange, as in reg_item: candidates
  // first (both sides in straight-line code, so their dependency chains
  // overlap), then committed for a running side only.
  struct Cand {
    bool conv, nonfin, brk;
    T alpha, p, s, sg;
  };
  auto cand = [&](double thr, T eta_new, T delta_new, T rr, T w_i, T u_i, T w_g, T p, T s_, T sg, T eta,
                  T alpha) -> Cand {
    Cand c;
    c.conv = it.rel ? ((double)rr <= thr) : (__builtin_fabs((double)eta_new) < thr);
    c.nonfin = !finite_t(eta_new);
    const T beta = qdiv(eta_new, eta);
    const T den = delta_new - beta * qdiv(eta_new, alpha);
    c.brk = den == (T)0 || !finite_t(den);
    c.alpha = uniform(qdiv(eta_new, den));
    c.p = fmaT(beta, p, u_i);
    c.s = fmaT(beta, s_, w_i);
    c.sg = fmaT(beta, sg, w_g);
    return c;
  };
  // returns whether the side keeps running
  auto commit = [&](const Cand &c, RegOut &out, uint32_t &iter, T eta_new, T rr, T &p, T &s_, T &sg, T &eta,
                    T &alpha, T (&last)[2]) -> bool {
    ++iter;
    last[0] = eta_new;
    last[1] = rr;
    if (c.conv) {
      out.status = GBDPCG_CONVERGED;
      return false;
    }
    if (c.nonfin) {
      out.status = GBDPCG_NONFINITE;
      return false;
    }
    if (c.brk) {
      out.status = GBDPCG_BREAKDOWN;
      return false;
    }
    alpha = c.alpha;
    eta = eta_new;
    p = c.p;
    s_ = c.s;
    if (it.ghost) sg = c.sg;
    return iter < max_iter;
  };
  // Both sides' vector work runs every iteration (a stopped side's results are
  // discarded and its state is frozen), so that the compiler can interleave the
  // two sides' matvecs, reductions and scalar chains.
  while (runA || runB) {
    if (it.owner) {
      const T xa = fmaT(alphaA, pA, xA), ra = fmaT(-alphaA, sA, rA_i);
      const T xb = fmaT(alphaB, pB, xB), rb = fmaT(-alphaB, sB, rB_i);
      if (runA) {
        xA = xa;
        rA_i = ra;
      }
      if (runB) {
        xB = xb;
        rB_i = rb;
      }
      rA[own_pos] = rA_i;
      rB[own_pos] = rB_i;
    }
    if (it.ghost) {
      const T ga = fmaT(-alphaA, sgA, rgA), gb = fmaT(-alphaB, sgB, rgB);
      if (runA) rgA = ga;
      if (runB) rgB = gb;
      rA[it.gpos] = rgA;
      rB[it.gpos] = rgB;
    }
    __syncthreads();
    if constexpr (PLDS) {
      pinv_mv2<PLDS>(it, mM, plds, rA, rB, t, uA_i, uB_i);
    } else {
      uA_i = row_mv(it, mM, rA, it.act_u);
      uB_i = row_mv(it, mM, rB, it.act_u);
    }
    if (it.act_u && c == 0) {
      uA[(m + 1) * nP + row] = uA_i;
      uB[(m + 1) * nP + row] = uB_i;
    }
    __syncthreads();
    if constexpr (SLDS) {
      s_mv2<SLDS, SL>(it, mS, slds, uA, uB, t, wA_i, wB_i);
    } else {
      wA_i = row_mv(it, mS, uA, it.act_w);
      wB_i = row_mv(it, mS, uB, it.act_w);
    }
    // both edge blocks every exchange (a stopped side's are stale and unused):
    // the neighbours' gather waits for every slot of the current epoch
    it.publish_edge(wA_i, 2);
    it.publish_edge(wB_i, 2, VB);
    it.reduce_publish6(it.owner ? rA_i * uA_i : (T)0, it.owner ? wA_i * uA_i : (T)0, it.owner ? rA_i * rA_i : (T)0,
                       it.owner ? rB_i * uB_i : (T)0, it.owner ? wB_i * uB_i : (T)0, it.owner ? rB_i * rB_i : (T)0);
    if (!it.gather_pair(wA, wB, 2, tt)) {
      timed_out = true;
      break;
    }
    const T wgA = it.ghost ? wA[it.gpos] : (T)0;
    const T wgB = it.ghost ? wB[it.gpos] : (T)0;
    const Cand cA = cand(thrA, tt[0], tt[1], tt[2], wA_i, uA_i, wgA, pA, sA, sgA, etaA, alphaA);
    const Cand cB = cand(thrB, tt[3], tt[4], tt[5], wB_i, uB_i, wgB, pB, sB, sgB, etaB, alphaB);
    if (runA) runA = commit(cA, fo, iterA, tt[0], tt[2], pA, sA, sgA, etaA, alphaA, lastA);
    if (runB) runB = commit(cB, ao, iterB, tt[3], tt[5], pB, sB, sgB, etaB, alphaB, lastB);
  }
  fo.iters = (int)iterA;
  ao.iters = (int)iterB;
  fo.eta = (double)lastA[0];
  fo.rr = (double)lastA[1];
  ao.eta = (double)lastB[0];
  ao.rr = (double)lastB[1];
  if (timed_out) {
    fo.status = GBDPCG_TIMED_OUT;
    ao.status = GBDPCG_TIMED_OUT;
  } else {
    // lambda and g_gamma to HBM and, with one ghost knot each side, to LDS
    if (it.owner) {
      gptr(static_cast<T *>(d.lambda))[(size_t)kg * n + row] = xA;
      gptr(static_cast<T *>(d.ggamma))[(size_t)kg * n + row] = xB;
      it.lv[own_pos] = xA;
      it.xv[own_pos] = xB;
    }
    it.publish_edge(xB, 1);
    it.publish_edge(xA, 1, VB);
    it.reduce_publish6((T)0, (T)0, (T)0, (T)0, (T)0, (T)0);
    if (!it.gather_pair(it.xv, it.lv, 1, tt)) {
      timed_out = true;
      fo.status = GBDPCG_TIMED_OUT;
      ao.status = GBDPCG_TIMED_OUT;
    } else {
      write_gS<T, NS, NT>(static_cast<T *>(d.gS), it.xv, it.lv, it.a, K, it.N);
    }
  }
  __syncthreads();
  if (rank == 0 && t == 0 && d.result != nullptr) {
    gbdpcg_result res;
    res.iters = fo.iters;
    res.status = fo.status;
    res.adj_iters = ao.iters;
    res.adj_status = ao.status;
    res.eta = fo.eta;
    res.rr = fo.rr;
    res.adj_eta = ao.eta;
    res.adj_rr = ao.rr;
    *d.result = res;
  }
  if (timed_out && t == 0 && d.err != nullptr)
    __hip_atomic_store(d.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace

// NS > 0: the kernel is specialised for stateSize NS.  Runtime dispatch for
// batches of uneven stateSize: NS == 0 over every specialised size, NS == -1
// over those up to 14 (plans that give rows of n = 28 a launch of their own;
// the instance then needs the registers of n <= 14 only).
template <typename T, int NS, int NT, bool FUSED, bool BC>
__device__ __forceinline__ void reg_dispatch(const KSys &d, uint32_t rank, T *lds) {
  if constexpr (NS > 0) {
    reg_item<T, NS, NT, FUSED, BC>(d, rank, lds);
  } else if constexpr (BC) {
    switch (d.n) {
      case 12: reg_item<T, 12, NT, FUSED, true, true, false>(d, rank, lds); break;
      case 14: reg_item<T, 14, NT, FUSED, true, true, false>(d, rank, lds); break;
      default: break;  // the planner only admits the sizes above
    }
  } else {
    switch (d.n) {
      case 2: reg_item<T, 2, NT, FUSED, false, true, false>(d, rank, lds); break;
      case 6: reg_item<T, 6, NT, FUSED, false, true, false>(d, rank, lds); break;
      case 12: reg_item<T, 12, NT, FUSED, false, true, false>(d, rank, lds); break;
      case 14: reg_item<T, 14, NT, FUSED, false, true, false>(d, rank, lds); break;
      case 28:
        if constexpr (NS == 0) reg_item<T, 28, NT, FUSED, false, true, false>(d, rank, lds);
        break;
      default: break;  // the planner only admits the sizes above
    }
  }
}

// Single system (systems == nullptr): workgroup b is rank b/stride when
// b % stride == 0.  Batched: items[round * gridDim.x + b] as in pcg_kernel.
template <typename T, int NS, int NT, bool FUSED, bool BC>
__device__ __forceinline__ void reg_entry(const KSys &one, const KSys *__restrict__ systems,
                                          const uint32_t *__restrict__ items, uint32_t rounds, uint32_t stride) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  if (systems == nullptr) {
    if (blockIdx.x % stride) return;
    const uint32_t rank = blockIdx.x / stride;
    if (rank >= one.G) return;
    reg_dispatch<T, NS, NT, FUSED, BC>(one, rank, lds);
    return;
  }
  for (uint32_t rd = 0; rd < rounds; ++rd) {
    const uint32_t item = items[(size_t)rd * gridDim.x + blockIdx.x];
    if (item == kNoItem) continue;
    __syncthreads();
    const KSys d = systems[item >> 12];
    reg_dispatch<T, NS, NT, FUSED, BC>(d, item & 0xFFFu, lds);
  }
}

// Waves per SIMD the column-chunk kernel is compiled for: the FP32 instances
// for rows of n <= 14 at 512 threads (the specialised ones and the runtime-n
// one, NS = -1) are held to 128 VGPRs, so that two workgroups share a CU in
// batch plans and long single solves (the planners count on it).
template <typename T, int NS, int NT>
constexpr int reg_waves_per_eu() {
  return (sizeof(T) == 4 && NT == 512 && NS != 0 && NS <= 14) ? 4 : 1;
}

// Register-resident rows, reg_lpr(n) lanes per row, column chunks.
template <typename T, int NS, int NT, bool FUSED>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(reg_waves_per_eu<T, NS, NT>())))
pcg_reg_kernel(KSys one, const KSys *__restrict__ systems,
                                                     const uint32_t *__restrict__ items, uint32_t rounds,
                                                     uint32_t stride) {
  reg_entry<T, NS, NT, FUSED, false>(one, systems, items, rounds, stride);
}

// Register-resident rows, row-broadcast layout (n in {12, 14}): one lane per row,
// 16 lanes per knot, the vector window broadcast within each 16-lane row by DPP.
template <typename T, int NS, int NT, bool FUSED>
__global__ void __launch_bounds__(NT) pcg_rowcast_kernel(KSys one, const KSys *__restrict__ systems,
                                                         const uint32_t *__restrict__ items, uint32_t rounds,
                                                         uint32_t stride) {
  reg_entry<T, NS, NT, FUSED, true>(one, systems, items, rounds, stride);
}

// Paired forward + adjoint items (fused recurrence, both flags set) go to
// reg_item_pair, every other item to reg_item.
template <typename T, int NS, int NT, bool BC>
__device__ __forceinline__ void pair_item(const KSys &d, uint32_t rank, T *lds) {
  if ((d.flags & (KF_FWD | KF_ADJ)) == (KF_FWD | KF_ADJ) && (d.flags & KF_FUSED)) {
    reg_item_pair<T, NS, NT, BC>(d, rank, lds);
  } else if constexpr (pair_s_lds(sizeof(T), NT, BC, NS)) {
    // instances with S in LDS run paired items only (plans route no other item
    // to them): flag the launch rather than skip the item silently
    if (threadIdx.x == 0 && d.err != nullptr) __hip_atomic_store(d.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  } else {
    reg_item<T, NS, NT, true, BC, false>(d, rank, lds);
  }
}

template <typename T, int NS, int NT, bool BC>
__device__ __forceinline__ void pair_dispatch(const KSys &d, uint32_t rank, T *lds) {
  if constexpr (NS > 0) {
    pair_item<T, NS, NT, BC>(d, rank, lds);
  } else if constexpr (BC) {
    switch (d.n) {
      case 12: pair_item<T, 12, NT, true>(d, rank, lds); break;
      case 14: pair_item<T, 14, NT, true>(d, rank, lds); break;
      default: break;
    }
  } else {
    switch (d.n) {  // runtime-n paired plans hold rows of n <= 14 only
      case 2: pair_item<T, 2, NT, false>(d, rank, lds); break;
      case 6: pair_item<T, 6, NT, false>(d, rank, lds); break;
      case 12: pair_item<T, 12, NT, false>(d, rank, lds); break;
      case 14: pair_item<T, 14, NT, false>(d, rank, lds); break;
      default: break;
    }
  }
}

// Waves per SIMD the paired kernel is compiled for: FP32 column-chunk rows of
// n <= 14 (and the runtime-n instance, which holds those sizes only) at 512
// threads are held to 128 VGPRs, so that two workgroups share a
// CU (a batch of 64 systems of N = 128 then runs in one round); the compiler
// spills what does not fit.
template <typename T, int NS, int NT, bool BC>
constexpr int pair_waves_per_eu() {
  return (sizeof(T) == 4 && NT == 512 && !BC && NS <= 14) ? 4 : 1;
}

// Forward + adjoint solves iterated together (see reg_item_pair); same launch
// interface as pcg_reg_kernel.
template <typename T, int NS, int NT, bool BC>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(pair_waves_per_eu<T, NS, NT, BC>())))
pcg_pair_kernel(KSys one, const KSys *__restrict__ systems,
                                                      const uint32_t *__restrict__ items, uint32_t rounds,
                                                      uint32_t stride) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  if (systems == nullptr) {
    if (blockIdx.x % stride) return;
    const uint32_t rank = blockIdx.x / stride;
    if (rank >= one.G) return;
    pair_dispatch<T, NS, NT, BC>(one, rank, lds);
    return;
  }
  for (uint32_t rd = 0; rd < rounds; ++rd) {
    const uint32_t item = items[(size_t)rd * gridDim.x + blockIdx.x];
    if (item == kNoItem) continue;
    __syncthreads();
    const KSys d = systems[item >> 12];
    pair_dispatch<T, NS, NT, BC>(d, item & 0xFFFu, lds);
  }
}

}  // namespace gbdpcg
