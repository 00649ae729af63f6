// Register-resident PCG kernel for gfx950: the latency path (templates; the
// fused-recurrence instances are compiled in pcg_reg_kernel.hip, the
// reference-recurrence ones in pcg_reg_ref_kernel.hip).
//
// Same method and exchange protocol as pcg_kernel.hip, with every block row a
// workgroup owns held in registers instead of LDS.  FUSED: the single-reduction
// (Chronopoulos-Gear) form, described below; reference form
// (include/pcg.cuh:118-208): two exchanges per iteration, the first carrying
// p'v and the edge knots of v = S p, the second r'z and r'r only (z = Pinv r is
// formed on one ghost knot each side from the resident ghost rows of Pinv, so
// p needs no halo).  Fused form:
//   * LPR lanes (reg_lpr: 1 for n <= 2, else 2) share one matrix row; each lane
//     keeps a CW-column chunk of that row of Pinv (rows of the owned knots plus
//     one ghost knot each side) and of S (owned knots) in VGPRs for the whole
//     solve, loaded from HBM once;
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

// Status of one exchange within the item (all threads agree after a gather).
template <typename T, int NS, int NT>
struct RegItem {
  static constexpr int n = NS;
  static constexpr int nP = (int)padded_n(NS);
  static constexpr int LPR = (int)reg_lpr(NS, NT, (int)sizeof(T));  // lanes per matrix row
  static constexpr int CW = (int)reg_chunk(NS, LPR);                 // columns per lane
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

  template <bool XCD>
  __device__ __forceinline__ void store_edges(uint64_t *r, T val, int edge, uint32_t epoch) const {
    const int lim = edge << 16;
    if (slot_lo >= 0 && slot_lo < lim) {
#pragma unroll
      for (int qq = 0; qq < WPV; ++qq) {
        uint64_t *p = r + (slot_lo & 0xFFFF) * WPV + qq;
        if constexpr (XCD) xstore_xcd(p, payload_word(val, qq), epoch); else xstore(p, payload_word(val, qq), epoch);
      }
    }
    if (slot_hi >= 0 && slot_hi < lim) {
#pragma unroll
      for (int qq = 0; qq < WPV; ++qq) {
        uint64_t *p = r + (slot_hi & 0xFFFF) * WPV + qq;
        if constexpr (XCD) xstore_xcd(p, payload_word(val, qq), epoch); else xstore(p, payload_word(val, qq), epoch);
      }
    }
  }

  __device__ __forceinline__ void publish_edge(T val, int edge) const {
    if (G == 1) return;
    const uint32_t epoch = base + seq;
    uint64_t *r = rec((int)(seq & 1u), rank);
    if (xcd)
      store_edges<true>(r, val, edge, epoch);
    else
      store_edges<false>(r, val, edge, epoch);
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
    T tv = wsum[l];
#pragma unroll
    for (int i = 1; i < NWAVES; ++i) tv += wsum[i * 3 + l];
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

  // Polls the current exchange.  Wave 0 gathers every rank's wave totals and
  // reduces them (waves in order, then ranks by a fixed DPP tree) into tot; the
  // other waves copy the neighbours' `edge` edge knots into the ghost knots of
  // `vec`.  Ends with a barrier after which t0..t2 hold the totals in every
  // thread; returns false (uniformly) if a wait hit its bound.
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
    if (t < 64) {
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
      acc0 = wave_sum(acc0);
      acc1 = wave_sum(acc1);
      acc2 = wave_sum(acc2);
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
        if (i < nleft) {  // left neighbour's right edge: knot a-1-e = local 1-e
          const int vi = i / WPV, qq = i - vi * WPV;
          const int e = vi / n, rr = vi - e * n;
          src = rec(type, rank - 1) + (kRecScalars + 2 * n + e * n + rr) * WPV + qq;
          dst = reinterpret_cast<uint32_t *>(vec + (1 - e) * nP + rr) + qq;
        } else {  // right neighbour's left edge: knot b+e = local K+2+e
          const int ii = i - nleft;
          const int vi = ii / WPV, qq = ii - vi * WPV;
          const int e = vi / n, rr = vi - e * n;
          src = rec(type, rank + 1) + (kRecScalars + e * n + rr) * WPV + qq;
          dst = reinterpret_cast<uint32_t *>(vec + (K + 2 + e) * nP + rr) + qq;
        }
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
template <typename T, int NS, int CW>
__device__ __forceinline__ void load_chunk(T (&dst)[CW], const T *__restrict__ A, int kg, int row, int c, int N,
                                           bool active) {
  constexpr int n = NS, nP = (int)padded_n(NS);
  // opaque copies: keeps the per-lane offsets below from being hoisted out of the
  // batched item loop (they would stay live, in registers, for the whole solve)
  int cl = c, rl = row;
  asm volatile("" : "+v"(cl), "+v"(rl));
#pragma unroll
  for (int j = 0; j < CW; ++j) {
    const int col = cl * CW + j;
    const int seg = col / nP, cc = col - seg * nP;
    T v = 0;
    if (active && kg >= 0 && kg < N && seg < 3 && cc < n && !(kg == 0 && seg == 0) && !(kg == N - 1 && seg == 2))
      v = gptr(A)[(size_t)kg * 3 * n * n + (size_t)(seg * n + cc) * n + rl];
    dst[j] = v;
  }
}

__device__ __forceinline__ bool finite_d(double v) { return __builtin_isfinite(v); }
__device__ __forceinline__ bool finite_t(float v) { return __builtin_isfinite(v); }
__device__ __forceinline__ bool finite_t(double v) { return __builtin_isfinite(v); }


template <typename T, int NS, int NT, bool FUSED>
__device__ __forceinline__ void reg_item(const KSys &d, uint32_t rank, T *lds) {
  using R = RegItem<T, NS, NT>;
  constexpr int n = R::n, nP = R::nP, CW = R::CW;
  R it;
  it.N = (int)d.N;
  it.G = (int)d.G;
  it.rank = (int)rank;
  it.a = (int)knot_begin(rank, d.G, d.N);
  it.b = (int)knot_begin(rank + 1, d.G, d.N);
  it.K = it.b - it.a;
  const int K = it.K;
  it.VL = (K + 4) * nP + R::LPR * CW;
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
  it.m = it.q / n;
  it.row = it.q - it.m * n;
  it.kg = it.a - 1 + it.m;
  it.act_u = it.m < K + 2;
  it.act_w = it.m >= 1 && it.m <= K;
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

  // matrix chunks straight from HBM into registers
  T mM[CW], mS[CW];
  load_chunk<T, NS, CW>(mM, static_cast<const T *>(d.Pinv), kg, row, c, it.N, it.act_u);
  load_chunk<T, NS, CW>(mS, static_cast<const T *>(d.S), kg, row, c, it.N, it.act_w);
  __syncthreads();
  it.base = (uint32_t)it.ctl[1];
  it.seq = 0;

  const bool fwd = (d.flags & KF_FWD) != 0;
  const bool adj = (d.flags & KF_ADJ) != 0;
  // initial guess window (own + one ghost knot each side)
  {
    const T *lam0 = static_cast<const T *>(d.lambda);
    T *dst = fwd ? it.xv : it.lv;
    for (int i = opaque_tid(); i < (K + 2) * n; i += NT) {
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
    const T *rhs = static_cast<const T *>(pass == 0 ? d.gamma : d.glambda);
    T x_i = 0, b_i = 0;
    if (it.owner) {
      x_i = it.xv[own_pos];
      b_i = gptr(rhs)[(size_t)kg * n + row];
    }
    // r = rhs - S x (own rows)
    T r_i;
    {
      T y = it.act_w ? chunk_dot<T, CW>(mS, it.xv + m * nP + c * CW) : (T)0;
      y = group_sum<R::LPR>(y);
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
        T y = it.act_u ? chunk_dot<T, CW>(mM, it.rv + m * nP + c * CW) : (T)0;
        z_i = group_sum<R::LPR>(y);
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
            T y = it.act_w ? chunk_dot<T, CW>(mS, it.uv + m * nP + c * CW) : (T)0;
            v_i = group_sum<R::LPR>(y);
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
            T y = it.act_u ? chunk_dot<T, CW>(mM, it.rv + m * nP + c * CW) : (T)0;
            z_i = group_sum<R::LPR>(y);
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
        T y = it.act_u ? chunk_dot<T, CW>(mM, it.rv + m * nP + c * CW) : (T)0;
        u_i = group_sum<R::LPR>(y);
        if (it.act_u && c == 0) it.uv[(m + 1) * nP + row] = u_i;
        __syncthreads();
        T z = it.act_w ? chunk_dot<T, CW>(mS, it.uv + m * nP + c * CW) : (T)0;
        w_i = group_sum<R::LPR>(z);
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
          T y = it.act_u ? chunk_dot<T, CW>(mM, it.rv + m * nP + c * CW) : (T)0;
          u_i = group_sum<R::LPR>(y);
          if (it.act_u && c == 0) it.uv[(m + 1) * nP + row] = u_i;
          __syncthreads();
          T z = it.act_w ? chunk_dot<T, CW>(mS, it.uv + m * nP + c * CW) : (T)0;
          w_i = group_sum<R::LPR>(z);
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
          // beta and eta'/alpha are independent: two divisions in flight at once
          const T beta = qdiv(eta_new, eta);
          const T den = delta_new - beta * qdiv(eta_new, alpha);
          if (den == (T)0 || !finite_t(den)) {
            out.status = GBDPCG_BREAKDOWN;
            break;
          }
          alpha = qdiv(eta_new, den);
          eta = eta_new;
          p_i = fmaT(beta, p_i, u_i);
          s_i = fmaT(beta, s_i, w_i);
          if (it.ghost) s_g = fmaT(beta, s_g, w_g);
        }
        out.iters = (int)iter;
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
    __hip_atomic_fetch_or(d.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

// NS != 0: the kernel is specialised for stateSize NS.  NS == 0: runtime
// dispatch over the specialised sizes (batches of uneven stateSize).
template <typename T, int NS, int NT, bool FUSED>
__device__ __forceinline__ void reg_dispatch(const KSys &d, uint32_t rank, T *lds) {
  if constexpr (NS != 0) {
    reg_item<T, NS, NT, FUSED>(d, rank, lds);
  } else {
    switch (d.n) {
      case 2: reg_item<T, 2, NT, FUSED>(d, rank, lds); break;
      case 6: reg_item<T, 6, NT, FUSED>(d, rank, lds); break;
      case 12: reg_item<T, 12, NT, FUSED>(d, rank, lds); break;
      case 14: reg_item<T, 14, NT, FUSED>(d, rank, lds); break;
      case 28: reg_item<T, 28, NT, FUSED>(d, rank, lds); break;
      default: break;  // the planner only admits the sizes above
    }
  }
}

// Single system (systems == nullptr): workgroup b is rank b/stride when
// b % stride == 0.  Batched: items[round * gridDim.x + b] as in pcg_kernel.
template <typename T, int NS, int NT, bool FUSED>
__global__ void __launch_bounds__(NT) pcg_reg_kernel(KSys one, const KSys *__restrict__ systems,
                                                     const uint32_t *__restrict__ items, uint32_t rounds,
                                                     uint32_t stride) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  if (systems == nullptr) {
    if (blockIdx.x % stride) return;
    const uint32_t rank = blockIdx.x / stride;
    if (rank >= one.G) return;
    reg_dispatch<T, NS, NT, FUSED>(one, rank, lds);
    return;
  }
  for (uint32_t rd = 0; rd < rounds; ++rd) {
    const uint32_t item = items[(size_t)rd * gridDim.x + blockIdx.x];
    if (item == kNoItem) continue;
    __syncthreads();
    const KSys d = systems[item >> 12];
    reg_dispatch<T, NS, NT, FUSED>(d, item & 0xFFFu, lds);
  }
}

}  // namespace gbdpcg
