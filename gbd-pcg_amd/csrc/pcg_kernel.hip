// Persistent block-tridiagonal PCG kernel for gfx950 (MI355X).
//
// One launch solves one system (drop-in path) or many systems of uneven size
// (batched path).  A system of N knots is split into G contiguous knot ranges;
// each range is owned by one workgroup ("rank").  A rank keeps its block rows
// of S and Pinv in LDS for the whole solve (loaded from HBM once, coalesced),
// and its slice of every iteration vector in LDS with ghost knots on both sides.
//
// Cross-workgroup communication happens only at the scalar reductions of the
// method.  Each reduction is an all-gather of tiny records through 64-bit words
// {32-bit payload, 32-bit epoch} written and polled with device-scope relaxed
// atomics: a reader spins until the word carries the epoch it expects, so no
// barrier, fence or flag/data ordering is needed.  The halo values a rank needs
// from its neighbours (edge knots of the vector just produced by a matvec) ride
// in the same records, so an iteration of the reference recurrence costs exactly
// two exchanges and the fused (Chronopoulos-Gear) recurrence one.
//
// The iteration itself restates the reference kernel include/pcg.cuh:118-208
// (r = gamma - S lambda; z = Pinv r; p = z; eta = r'z; loop: v = S p,
// alpha = eta / p'v, lambda += alpha p, r -= alpha v, z = Pinv r, eta' = r'z,
// exit on |eta'| < tol with iters = i + 1, beta = eta'/eta, p = z + beta p).
// The block-row matvec is the reference's bdmv (include/utils.cuh:46-85) with
// the halo gather of loadbdVec (include/utils.cuh:9-40) replaced by ghost knots.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "device_common.h"
#include "internal.h"

namespace gbdpcg {

// Sum of three per-thread values over the workgroup; every thread gets the
// result.  Contains one __syncthreads.
template <typename T>
__device__ __forceinline__ void block_sum3(T &a, T &b, T &c, T *scratch) {
  wave_sum3(a, b, c);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    scratch[w * 3 + 0] = a;
    scratch[w * 3 + 1] = b;
    scratch[w * 3 + 2] = c;
  }
  __syncthreads();
  a = ((scratch[0] + scratch[3]) + scratch[6]) + scratch[9];
  b = ((scratch[1] + scratch[4]) + scratch[7]) + scratch[10];
  c = ((scratch[2] + scratch[5]) + scratch[8]) + scratch[11];
}

// Sum over the G ranks of scalar s from the gathered partials part[g*3 + s].
template <typename T>
__device__ __forceinline__ T sum_ranks(const T *part, int G, int s) {
  const int l = threadIdx.x & 63;
  T acc = 0;
  for (int g = l; g < G; g += 64) acc += part[g * 3 + s];
  return wave_sum(acc);
}

// ---------------------------------------------------------------------------
// on-chip block-row matvec
// ---------------------------------------------------------------------------

// Dot product of one CW-column chunk of a padded LDS row with the matching
// window of a ghost-extended vector.  CW is a multiple of 4 (16-byte LDS reads).
template <typename T, int CWc>
__device__ __forceinline__ T dot_chunk(const T *__restrict__ m, const T *__restrict__ x, int cw_rt) {
  const int cw = CWc ? CWc : cw_rt;
  T a0 = 0, a1 = 0;
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int j = 0; j < cw; j += 4) {
      const float4 mv = *reinterpret_cast<const float4 *>(m + j);
      const float4 xv = *reinterpret_cast<const float4 *>(x + j);
      a0 = __builtin_fmaf(mv.x, xv.x, a0);
      a1 = __builtin_fmaf(mv.y, xv.y, a1);
      a0 = __builtin_fmaf(mv.z, xv.z, a0);
      a1 = __builtin_fmaf(mv.w, xv.w, a1);
    }
  } else {
#pragma unroll
    for (int j = 0; j < cw; j += 4) {
      const double2 m0 = *reinterpret_cast<const double2 *>(m + j);
      const double2 m1 = *reinterpret_cast<const double2 *>(m + j + 2);
      const double2 x0 = *reinterpret_cast<const double2 *>(x + j);
      const double2 x1 = *reinterpret_cast<const double2 *>(x + j + 2);
      a0 = __builtin_fma(m0.x, x0.x, a0);
      a1 = __builtin_fma(m0.y, x0.y, a1);
      a0 = __builtin_fma(m1.x, x1.x, a0);
      a1 = __builtin_fma(m1.y, x1.y, a1);
    }
  }
  return a0 + a1;
}

// Geometry of one workgroup's share of a system.
struct Geo {
  int n, nP, RS, CW, N, G, rank, a, b, K, VL;
};

// y = M x for `R` consecutive knot rows stored in LDS at M (row-major, stride RS,
// n rows per knot).  Knot row m reads the 3 knots of x starting at local knot
// (m + wo) and the epilogue receives the local knot index of the output
// (m + oo), the row within the knot and the value.  kColSplit lanes share one
// output row and combine their chunks with xor shuffles.
template <typename T, int NS, typename Epi>
__device__ __forceinline__ void matvec(const Geo &g, const T *__restrict__ M, int R, const T *__restrict__ x,
                                       int wo, int oo, Epi &&epi) {
  constexpr int CWc = NS ? (int)col_chunk(NS) : 0;
  const int n = NS ? NS : g.n;
  const int units = R * n;
  const int lanes = units * kColSplit;
  for (int base = 0; base < lanes; base += kThreads) {
    const int L = base + (int)threadIdx.x;
    const int u = L / kColSplit;
    const int c = L % kColSplit;
    T y = 0;
    int m = 0, row = 0;
    if (u < units) {
      m = u / n;
      row = u - m * n;
      const T *mrow = M + (size_t)u * g.RS + c * g.CW;
      const T *xw = x + (m + wo) * g.nP + c * g.CW;
      y = dot_chunk<T, CWc>(mrow, xw, g.CW);
    }
    static_assert(kColSplit == 4, "the chunk reduction below is a quad reduction");
    y = quad_sum(y);
    if (u < units && c == 0) epi(m + oo, row, y);
  }
}

// ---------------------------------------------------------------------------
// loading cbtd block rows into padded row-major LDS rows
// ---------------------------------------------------------------------------

// Copies global block rows [k0, k1) of a cbtd matrix into LDS rows starting at
// local knot row `r0`.  The (0,-1) and (N-1,N) padding blocks are not read
// (the reference skips them too, include/pcg.cuh:105-106); the LDS already
// holds zeros there.
template <typename T, int NS>
__device__ __forceinline__ void load_rows(const Geo &g, T *__restrict__ dst, const T *__restrict__ src, int k0, int k1, int r0) {
  if (k1 <= k0) return;
  const int n = NS ? NS : g.n;
  const int bs = 3 * n * n;
  const size_t e0 = (size_t)k0 * bs;
  const int total = (k1 - k0) * bs;
  const T *base = src + e0;
  auto put = [&](int rel, T val) {
    const int kl = rel / bs;
    const int w = rel - kl * bs;
    const int col = w / n;
    const int rr = w - col * n;
    const int seg = col / n;
    const int cc = col - seg * n;
    const int k = k0 + kl;
    if ((k == 0 && seg == 0) || (k == g.N - 1 && seg == 2)) return;
    dst[(size_t)((r0 + kl) * n + rr) * g.RS + seg * g.nP + cc] = val;
  };
  constexpr int V = 16 / sizeof(T);
  if ((((uintptr_t)base) & 15) == 0 && (bs % V) == 0) {
    typedef T VT __attribute__((ext_vector_type(16 / sizeof(T))));
    const VT *vb = reinterpret_cast<const VT *>(base);
    for (int i = threadIdx.x; i < total / V; i += kThreads) {
      const VT v = gptr(vb)[i];
      const T *vv = reinterpret_cast<const T *>(&v);
#pragma unroll
      for (int q = 0; q < V; ++q) put(i * V + q, vv[q]);
    }
  } else {
    for (int i = threadIdx.x; i < total; i += kThreads) put(i, gptr(base)[i]);
  }
}

// Loads a knot-major global vector into a ghost-extended LDS vector: local knot
// j holds global knot a-2+j (zeros outside [0,N) and in the padding lanes).
// Only local knots [j0, j1) are filled.
template <typename T>
__device__ __forceinline__ void load_vec(const Geo &g, T *__restrict__ dst, const T *__restrict__ src, int j0, int j1) {
  const int cnt = (j1 - j0) * g.n;
  for (int i = threadIdx.x; i < cnt; i += kThreads) {
    const int jj = i / g.n;
    const int e = i - jj * g.n;
    const int j = j0 + jj;
    const int k = g.a - 2 + j;
    if (k >= 0 && k < g.N) dst[j * g.nP + e] = gptr(src)[(size_t)k * g.n + e];
  }
}

// ---------------------------------------------------------------------------
// per-item context
// ---------------------------------------------------------------------------

template <typename T>
struct Ctx {
  Geo g;
  uint64_t *slots;  // exchange records of the system
  double tol;
  uint32_t max_iter;
  T *Sm;      // K knot rows of S
  T *Pm;      // K+2 knot rows of Pinv (ghost rows first and last)
  T *x;       // solution (lambda or g_gamma)
  T *r;       // residual
  T *p;       // search direction
  T *z;       // preconditioned residual (u)
  T *v;       // S p (s in the fused form); holds the right-hand side during init
  T *w;       // S u (fused form)
  T *lam;     // forward solution kept for the adjoint's g_S
  T *wsum;    // 12: per-wave partials
  T *part;    // 3*G: gathered per-rank partials
  int *ctl;   // [0] failure flag, [1] epoch base
  uint32_t base;   // first epoch of this item
  uint32_t seq;    // exchanges done so far in this item
  int edge;        // edge knots carried per side (1 reference, 2 fused)
  bool fused, rel;
  bool xcd;        // every rank on this workgroup's XCD: exchange stores stay in L2
  bool xcd_ok;     // the system allows XCD-local stores (KF_NOXCD clear)
};

// Publishes this rank's record for the next exchange: scalars s0..s2 and the
// `edge` first/last owned knots of vector `vec`.
template <typename T>
__device__ __forceinline__ void publish(Ctx<T> &c, T s0, T s1, T s2, const T *vec) {
  constexpr int WPV = sizeof(T) / 4;
  const Geo &g = c.g;
  const uint32_t epoch = c.base + c.seq;
  const int type = (int)(c.seq & 1u);
  uint64_t *rec = c.slots + (size_t)(type * g.G + g.rank) * rec_words(g.n, WPV);
  const int en = c.edge * g.n;
  const int used = (3 + 2 * en) * WPV;
  for (int t = threadIdx.x; t < used; t += kThreads) {
    const int vi = t / WPV;
    const int q = t - vi * WPV;
    T val;
    if (vi < 3) {
      val = vi == 0 ? s0 : (vi == 1 ? s1 : s2);
    } else if (vi < 3 + en) {  // left edge: own knots a, a+1  (local 2, 3)
      const int e = (vi - 3) / g.n, i = (vi - 3) - e * g.n;
      val = vec[(2 + e) * g.nP + i];
    } else {  // right edge: own knots b-1, b-2 (local K+1, K)
      const int e = (vi - 3 - en) / g.n, i = (vi - 3 - en) - e * g.n;
      val = vec[(g.K + 1 - e) * g.nP + i];
    }
    // record layout (value index -> words [vi*WPV, vi*WPV + WPV)): scalars at 0..2,
    // left edge at kRecScalars + e*n + i, right edge at kRecScalars + 2n + e*n + i
    const int slot_vi = vi < 3 ? vi : (vi < 3 + en ? kRecScalars + (vi - 3) : kRecScalars + 2 * g.n + (vi - 3 - en));
    if (c.xcd)
      xstore_xcd(rec + slot_vi * WPV + q, payload_word(val, q), epoch);
    else
      xstore(rec + slot_vi * WPV + q, payload_word(val, q), epoch);
  }
}

// Waits for every rank's record of the current exchange: gathers the scalar
// partials into c.part and the neighbours' edge knots into the ghost knots of
// `vec`.  Ends with a __syncthreads; returns false (uniformly) on a timeout.
template <typename T>
__device__ __forceinline__ bool gather(Ctx<T> &c, T *vec) {
  constexpr int WPV = sizeof(T) / 4;
  const Geo &g = c.g;
  const uint32_t epoch = c.base + c.seq;
  const int type = (int)(c.seq & 1u);
  const uint32_t rw = rec_words(g.n, WPV);
  uint64_t *recs = c.slots + (size_t)(type * g.G) * rw;
  const int en = c.edge * g.n;
  const int nscal = g.G * 3 * WPV;
  const int nleft = g.rank > 0 ? en * WPV : 0;
  const int nright = g.rank < g.G - 1 ? en * WPV : 0;
  const int total = nscal + nleft + nright;
  bool ok = true;
  for (int t = threadIdx.x; t < total; t += kThreads) {
    uint64_t *src;
    uint32_t *dst;
    if (t < nscal) {
      const int rk = t / (3 * WPV);
      const int rem = t - rk * 3 * WPV;
      src = recs + (size_t)rk * rw + rem;
      dst = reinterpret_cast<uint32_t *>(c.part + rk * 3) + rem;
    } else if (t < nscal + nleft) {
      // left neighbour's right edge: its knot b'-1-e = our a-1-e = local 1-e
      const int tt = t - nscal;
      const int vi = tt / WPV, q = tt - vi * WPV;
      const int e = vi / g.n, i = vi - e * g.n;
      src = recs + (size_t)(g.rank - 1) * rw + (kRecScalars + 2 * g.n + e * g.n + i) * WPV + q;
      dst = reinterpret_cast<uint32_t *>(vec + (1 - e) * g.nP + i) + q;
    } else {
      // right neighbour's left edge: its knot a'+e = our b+e = local K+2+e
      const int tt = t - nscal - nleft;
      const int vi = tt / WPV, q = tt - vi * WPV;
      const int e = vi / g.n, i = vi - e * g.n;
      src = recs + (size_t)(g.rank + 1) * rw + (kRecScalars + e * g.n + i) * WPV + q;
      dst = reinterpret_cast<uint32_t *>(vec + (g.K + 2 + e) * g.nP + i) + q;
    }
    uint32_t val = 0;
    if (xwait(src, epoch, val)) {
      *dst = val;
    } else {
      ok = false;
      break;  // one failed wait is enough: do not spend the bound again
    }
  }
  if (!ok) c.ctl[0] = 1;
  __syncthreads();
  c.seq++;
  return c.ctl[0] == 0;
}

// Elementwise update over the ghost-extended range of local knots [j0, j1).
template <typename T, typename F>
__device__ __forceinline__ void for_range(const Ctx<T> &c, int j0, int j1, F &&f) {
  const int lo = j0 * c.g.nP, hi = j1 * c.g.nP;
  for (int i = lo + (int)threadIdx.x; i < hi; i += kThreads) f(i);
}

__device__ __forceinline__ bool finite_d(double v) { return __builtin_isfinite(v); }

template <typename T>
__device__ __forceinline__ bool exit_test(const Ctx<T> &c, T eta, T rr, T gg) {
  if (c.rel) return (double)rr <= c.tol * c.tol * (double)gg;
  return __builtin_fabs((double)eta) < c.tol;
}

struct CoreOut {
  int iters, status;
  double eta, rr;
};

// One PCG solve S x = rhs on the resident matrices.  x holds the initial guess
// (ghost-extended, width 2); rhs is a global knot-major vector.  Returns with x
// the solution on own knots and on one ghost knot each side.
template <typename T, int NS>
__device__ __forceinline__ CoreOut pcg_core(Ctx<T> &c, const T *__restrict__ rhs) {
  const Geo &g = c.g;
  const int K = g.K;
  CoreOut out{0, GBDPCG_MAX_ITER, 0.0, 0.0};
  T *x = c.x, *r = c.r, *p = c.p, *z = c.z, *v = c.v, *w = c.w;

  // right-hand side of the own knots into v (free until the first S p)
  load_vec(g, v, rhs, 2, K + 2);
  __syncthreads();

  // r = rhs - S x on own knots; gg = rhs'rhs
  T gg = 0, t1 = 0, t2 = 0;
  matvec<T, NS>(g, c.Sm, K, x, 1, 2, [&](int j, int row, T y) {
    const int i = j * g.nP + row;
    const T b = v[i];
    r[i] = b - y;
    gg += b * b;
  });
  block_sum3(gg, t1, t2, c.wsum);
  // the first exchange also carries x = 1 + this XCD's id and x^2: all ranks share
  // one XCD exactly when G * sum(x^2) == sum(x)^2 (small integers, exact in T)
  const T xid = (T)(1 + xcc_id());
  publish(c, gg, xid, xid * xid, r);
  if (!gather(c, r)) { out.status = GBDPCG_TIMED_OUT; return out; }
  gg = sum_ranks(c.part, g.G, 0);
  {
    const double sx = (double)sum_ranks(c.part, g.G, 1), sxx = (double)sum_ranks(c.part, g.G, 2);
    c.xcd = __builtin_amdgcn_readfirstlane((int)(c.xcd_ok && g.G > 1 && sx * sx == (double)g.G * sxx)) != 0;
  }

  T eta, rr, alpha = 0, beta = 0, delta = 0;
  if (!c.fused) {
    // z = Pinv r (own); eta = r'z
    T rz = 0, rr_ = 0, t = 0;
    matvec<T, NS>(g, c.Pm + (size_t)g.n * g.RS, K, r, 1, 2, [&](int j, int row, T y) {
      const int i = j * g.nP + row;
      z[i] = y;
      rz += r[i] * y;
      rr_ += r[i] * r[i];
    });
    block_sum3(rz, t, rr_, c.wsum);
    publish(c, rz, (T)0, rr_, z);
    if (!gather(c, z)) { out.status = GBDPCG_TIMED_OUT; return out; }
    eta = sum_ranks(c.part, g.G, 0);
    rr = sum_ranks(c.part, g.G, 2);
    for_range(c, 1, K + 3, [&](int i) { p[i] = z[i]; });
  } else {
    // u = Pinv r on own + one ghost knot each side; w = S u (own)
    T rz = 0, rr_ = 0, wu = 0;
    matvec<T, NS>(g, c.Pm, K + 2, r, 0, 1, [&](int j, int row, T y) {
      const int i = j * g.nP + row;
      z[i] = y;
      if (j >= 2 && j < K + 2) {
        rz += r[i] * y;
        rr_ += r[i] * r[i];
      }
    });
    __syncthreads();
    matvec<T, NS>(g, c.Sm, K, z, 1, 2, [&](int j, int row, T y) {
      const int i = j * g.nP + row;
      w[i] = y;
      wu += z[i] * y;
    });
    block_sum3(rz, wu, rr_, c.wsum);
    publish(c, rz, wu, rr_, w);
    if (!gather(c, w)) { out.status = GBDPCG_TIMED_OUT; return out; }
    eta = sum_ranks(c.part, g.G, 0);
    delta = sum_ranks(c.part, g.G, 1);
    rr = sum_ranks(c.part, g.G, 2);
    alpha = eta / delta;
    for_range(c, 1, K + 3, [&](int i) { p[i] = z[i]; });
    for_range(c, 0, K + 4, [&](int i) { v[i] = w[i]; });
  }
  out.eta = (double)eta;
  out.rr = (double)rr;
  if (!finite_d((double)eta) || !finite_d((double)rr)) {
    out.status = GBDPCG_NONFINITE;
    return out;
  }
  if (eta == (T)0) {  // r = 0 (converged) or r'Pinv r = 0 with r != 0 (breakdown)
    out.status = rr == (T)0 ? GBDPCG_CONVERGED : GBDPCG_BREAKDOWN;
    return out;
  }
  if (c.fused && (delta == (T)0 || !finite_d((double)alpha))) {
    out.status = GBDPCG_BREAKDOWN;
    return out;
  }
  __syncthreads();

  const uint32_t max_iter = c.max_iter;
  uint32_t iter = 0;
  for (; iter < max_iter;) {
    if (!c.fused) {
      // v = S p (own); alpha = eta / p'v
      T pv = 0, ta = 0, tb = 0;
      matvec<T, NS>(g, c.Sm, K, p, 1, 2, [&](int j, int row, T y) {
        const int i = j * g.nP + row;
        v[i] = y;
        pv += p[i] * y;
      });
      block_sum3(pv, ta, tb, c.wsum);
      publish(c, pv, (T)0, (T)0, v);
      if (!gather(c, v)) { out.status = GBDPCG_TIMED_OUT; out.iters = (int)iter; return out; }
      pv = sum_ranks(c.part, g.G, 0);
      if (pv == (T)0 || !finite_d((double)pv)) {
        out.status = pv == (T)0 ? GBDPCG_BREAKDOWN : GBDPCG_NONFINITE;
        break;
      }
      alpha = eta / pv;
      // lambda += alpha p ; r -= alpha v   (own knots and one ghost knot each side)
      for_range(c, 1, K + 3, [&](int i) {
        x[i] = fmaT(alpha, p[i], x[i]);
        r[i] = fmaT(-alpha, v[i], r[i]);
      });
      __syncthreads();
      // z = Pinv r ; eta' = r'z
      T rz = 0, rr_ = 0, tc = 0;
      matvec<T, NS>(g, c.Pm + (size_t)g.n * g.RS, K, r, 1, 2, [&](int j, int row, T y) {
        const int i = j * g.nP + row;
        z[i] = y;
        rz += r[i] * y;
        rr_ += r[i] * r[i];
      });
      block_sum3(rz, tc, rr_, c.wsum);
      publish(c, rz, (T)0, rr_, z);
      if (!gather(c, z)) { out.status = GBDPCG_TIMED_OUT; out.iters = (int)iter; return out; }
      const T eta_new = sum_ranks(c.part, g.G, 0);
      rr = sum_ranks(c.part, g.G, 2);
      ++iter;
      out.eta = (double)eta_new;
      out.rr = (double)rr;
      if (exit_test(c, eta_new, rr, gg)) { out.status = GBDPCG_CONVERGED; break; }
      if (!finite_d((double)eta_new)) { out.status = GBDPCG_NONFINITE; break; }
      beta = eta_new / eta;
      eta = eta_new;
      for_range(c, 1, K + 3, [&](int i) { p[i] = fmaT(beta, p[i], z[i]); });
      __syncthreads();
    } else {
      // x += alpha p (own + 1 ghost) ; r -= alpha s (own + 2 ghosts)
      for_range(c, 1, K + 3, [&](int i) { x[i] = fmaT(alpha, p[i], x[i]); });
      for_range(c, 0, K + 4, [&](int i) { r[i] = fmaT(-alpha, v[i], r[i]); });
      __syncthreads();
      T rz = 0, rr_ = 0, wu = 0;
      matvec<T, NS>(g, c.Pm, K + 2, r, 0, 1, [&](int j, int row, T y) {
        const int i = j * g.nP + row;
        z[i] = y;
        if (j >= 2 && j < K + 2) {
          rz += r[i] * y;
          rr_ += r[i] * r[i];
        }
      });
      __syncthreads();
      matvec<T, NS>(g, c.Sm, K, z, 1, 2, [&](int j, int row, T y) {
        const int i = j * g.nP + row;
        w[i] = y;
        wu += z[i] * y;
      });
      block_sum3(rz, wu, rr_, c.wsum);
      publish(c, rz, wu, rr_, w);
      if (!gather(c, w)) { out.status = GBDPCG_TIMED_OUT; out.iters = (int)iter; return out; }
      const T eta_new = sum_ranks(c.part, g.G, 0);
      const T delta_new = sum_ranks(c.part, g.G, 1);
      rr = sum_ranks(c.part, g.G, 2);
      ++iter;
      out.eta = (double)eta_new;
      out.rr = (double)rr;
      if (exit_test(c, eta_new, rr, gg)) { out.status = GBDPCG_CONVERGED; break; }
      if (!finite_d((double)eta_new)) { out.status = GBDPCG_NONFINITE; break; }
      beta = eta_new / eta;
      const T den = delta_new - beta * eta_new / alpha;
      if (den == (T)0 || !finite_d((double)den)) { out.status = GBDPCG_BREAKDOWN; break; }
      alpha = eta_new / den;
      eta = eta_new;
      // p = u + beta p (own + 1 ghost) ; s = w + beta s (own + 2 ghosts)
      for_range(c, 1, K + 3, [&](int i) { p[i] = fmaT(beta, p[i], z[i]); });
      for_range(c, 0, K + 4, [&](int i) { v[i] = fmaT(beta, v[i], w[i]); });
      __syncthreads();
    }
  }
  out.iters = (int)iter;
  return out;
}

// Writes own knots of a ghost-extended LDS vector to a global knot-major vector.
template <typename T>
__device__ __forceinline__ void store_vec(const Geo &g, T *__restrict__ dst, const T *__restrict__ src) {
  const int cnt = g.K * g.n;
  for (int i = threadIdx.x; i < cnt; i += kThreads) {
    const int jj = i / g.n;
    const int e = i - jj * g.n;
    gptr(dst)[(size_t)(g.a + jj) * g.n + e] = src[(jj + 2) * g.nP + e];
  }
}

// g_S block rows of the own knots: G_{k,j} = -1/2 (gg_k lam_j^T + lam_k gg_j^T)
// for j in {k-1, k, k+1} (zero outside the system), cbtd column-major.
template <typename T>
__device__ __forceinline__ void store_gS(const Geo &g, T *__restrict__ gS, const T *__restrict__ lam, const T *__restrict__ gg) {
  const int n = g.n;
  const int bs = 3 * n * n;
  const int cnt = g.K * bs;
  for (int e = threadIdx.x; e < cnt; e += kThreads) {
    const int kl = e / bs;
    const int wq = e - kl * bs;
    const int col = wq / n;
    const int rr = wq - col * n;
    const int seg = col / n;
    const int cc = col - seg * n;
    const int k = g.a + kl;
    const int j = k - 1 + seg;
    T val = 0;
    if (j >= 0 && j < g.N) {
      const int lk = (kl + 2) * g.nP, lj = (kl + 1 + seg) * g.nP;
      val = (T)-0.5 * (gg[lk + rr] * lam[lj + cc] + lam[lk + rr] * gg[lj + cc]);
    }
    gptr(gS)[(size_t)g.a * bs + e] = val;
  }
}

template <typename T, int NS>
__device__ __forceinline__ void run_item(const KSys &d, uint32_t rank, T *lds) {
  Ctx<T> c;
  Geo &g = c.g;
  g.n = NS ? NS : (int)d.n;
  g.nP = (int)padded_n(g.n);
  g.CW = (int)col_chunk(g.n);
  g.RS = (int)row_stride(g.n);
  g.N = (int)d.N;
  g.G = (int)d.G;
  g.rank = (int)rank;
  g.a = (int)knot_begin(rank, d.G, d.N);
  g.b = (int)knot_begin(rank + 1, d.G, d.N);
  g.K = g.b - g.a;
  g.VL = (g.K + 4) * g.nP + g.RS;
  c.slots = d.slots;
  c.tol = d.tol;
  c.max_iter = d.max_iter;
  c.fused = (d.flags & KF_FUSED) != 0;
  c.rel = (d.flags & KF_REL) != 0;
  c.edge = c.fused ? 2 : 1;
  c.xcd = false;
  c.xcd_ok = (d.flags & KF_NOXCD) == 0;
  const size_t mat = (size_t)g.K * g.n * g.RS;
  c.Sm = lds;
  c.Pm = c.Sm + mat;
  c.x = c.Pm + (size_t)(g.K + 2) * g.n * g.RS;
  c.r = c.x + g.VL;
  c.p = c.r + g.VL;
  c.z = c.p + g.VL;
  c.v = c.z + g.VL;
  c.w = c.v + g.VL;
  c.lam = c.w + g.VL;
  c.wsum = c.lam + g.VL;
  c.part = c.wsum + 12;
  c.ctl = reinterpret_cast<int *>(c.part + 3 * g.G);
  const size_t total = (size_t)(c.part + 3 * g.G - lds);

  // epoch base: one past the newest epoch in this rank's own records (all ranks
  // of the system hold the same two values, so all agree on the base)
  constexpr int WPV = sizeof(T) / 4;
  if (threadIdx.x == 0) {
    const uint32_t rw = rec_words(g.n, WPV);
    const uint32_t ea = (uint32_t)(xload(d.slots + (size_t)g.rank * rw) >> 32);
    const uint32_t eb = (uint32_t)(xload(d.slots + (size_t)(g.G + g.rank) * rw) >> 32);
    uint32_t base = (ea > eb ? ea : eb) + 1u;
    if (base > 0xFFF00000u) base = 1u;  // far from wrapping: restart the sequence
    c.ctl[0] = 0;
    c.ctl[1] = (int)base;
  }
  // zero the matrices and vectors (padding lanes, ghost knots, pad blocks)
  for (size_t i = threadIdx.x; i < total; i += kThreads) lds[i] = (T)0;
  __syncthreads();
  c.base = (uint32_t)c.ctl[1];
  c.seq = 0;

  const T *S = static_cast<const T *>(d.S);
  const T *P = static_cast<const T *>(d.Pinv);
  load_rows<T, NS>(g, c.Sm, S, g.a, g.b, 0);
  {
    const int k0 = g.a - 1 < 0 ? 0 : g.a - 1;
    const int k1 = g.b + 1 > g.N ? g.N : g.b + 1;
    load_rows<T, NS>(g, c.Pm, P, k0, k1, k0 - (g.a - 1));
  }
  const bool fwd = (d.flags & KF_FWD) != 0;
  const bool adj = (d.flags & KF_ADJ) != 0;
  load_vec(g, fwd ? c.x : c.lam, static_cast<const T *>(d.lambda), 0, g.K + 4);
  __syncthreads();

  CoreOut fo{0, GBDPCG_CONVERGED, 0.0, 0.0}, ao{0, GBDPCG_CONVERGED, 0.0, 0.0};
  bool timed_out = false;
  // pass 0: forward solve S lambda = gamma; pass 1: adjoint S g_gamma = g_lambda.
  // One call site keeps the solver body inlined once.
  for (int pass = fwd ? 0 : 1; pass <= (adj ? 1 : 0); ++pass) {
    if (pass == 1 && fwd) {
      for (int i = threadIdx.x; i < g.VL; i += kThreads) {
        c.lam[i] = c.x[i];
        c.x[i] = (T)0;
      }
      __syncthreads();
    }
    const CoreOut o = pcg_core<T, NS>(c, static_cast<const T *>(pass == 0 ? d.gamma : d.glambda));
    if (pass == 0) fo = o; else ao = o;
    if (o.status == GBDPCG_TIMED_OUT) {
      timed_out = true;
      break;
    }
    if (pass == 0) {
      store_vec(g, static_cast<T *>(d.lambda), c.x);
    } else {
      store_vec(g, static_cast<T *>(d.ggamma), c.x);
      store_gS(g, static_cast<T *>(d.gS), c.lam, c.x);
    }
    __syncthreads();
  }
  if (rank == 0 && threadIdx.x == 0 && d.result != nullptr) {
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
  // any rank that timed out flags the launch, so the host never misses a failure
  if (timed_out && threadIdx.x == 0 && d.err != nullptr) {
    __hip_atomic_store(d.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

template <typename T>
__device__ __forceinline__ void run_dispatch(const KSys &d, uint32_t rank, T *lds) {
  switch (d.n) {
    case 2: run_item<T, 2>(d, rank, lds); break;
    case 6: run_item<T, 6>(d, rank, lds); break;
    case 12: run_item<T, 12>(d, rank, lds); break;
    case 14: run_item<T, 14>(d, rank, lds); break;
    case 28: run_item<T, 28>(d, rank, lds); break;
    default: run_item<T, 0>(d, rank, lds); break;
  }
}

// Single system: `one` by value; workgroup b works as rank b/stride when
// b % stride == 0 (stride 8 keeps every rank on one XCD: workgroups are dealt
// to the 8 XCDs round-robin).
// Batched: items[round * gridDim.x + b] lists the (system, rank) pairs of
// workgroup b in an order consistent across workgroups (no cyclic waits).
template <typename T>
__global__ void __launch_bounds__(kThreads) pcg_kernel(KSys one, const KSys *__restrict__ systems,
                                                       const uint32_t *__restrict__ items, uint32_t rounds,
                                                       uint32_t stride) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  if (systems == nullptr) {
    if (blockIdx.x % stride) return;
    const uint32_t rank = blockIdx.x / stride;
    if (rank >= one.G) return;
    run_dispatch<T>(one, rank, lds);
    return;
  }
  for (uint32_t rd = 0; rd < rounds; ++rd) {
    const uint32_t it = items[(size_t)rd * gridDim.x + blockIdx.x];
    if (it == kNoItem) continue;
    __syncthreads();
    const KSys d = systems[it >> 12];
    run_dispatch<T>(d, it & 0xFFFu, lds);
  }
}

template __global__ void pcg_kernel<float>(KSys, const KSys *, const uint32_t *, uint32_t, uint32_t);
template __global__ void pcg_kernel<double>(KSys, const KSys *, const uint32_t *, uint32_t, uint32_t);

}  // namespace gbdpcg

// ---------------------------------------------------------------------------
// host-side launch helpers (declared in internal.h)
// ---------------------------------------------------------------------------
namespace gbdpcg {

template <typename T>
static const void *pcg_kernel_ptr() {
  return reinterpret_cast<const void *>(&pcg_kernel<T>);
}

hipError_t pcg_prepare(int dtype, size_t lds_bytes) {
  const void *k = dtype == GBDPCG_F64 ? pcg_kernel_ptr<double>() : pcg_kernel_ptr<float>();
  return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}

hipError_t pcg_occupancy(int dtype, size_t lds_bytes, int *wgs_per_cu) {
  const void *k = dtype == GBDPCG_F64 ? pcg_kernel_ptr<double>() : pcg_kernel_ptr<float>();
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(wgs_per_cu, k, kThreads, lds_bytes);
}

hipError_t pcg_launch(int dtype, const KSys &one, const KSys *systems, const uint32_t *items, uint32_t rounds,
                      uint32_t stride, uint32_t grid, size_t lds_bytes, hipStream_t stream) {
  if (dtype == GBDPCG_F64) {
    hipLaunchKernelGGL(pcg_kernel<double>, dim3(grid), dim3(kThreads), lds_bytes, stream, one, systems, items,
                       rounds, stride);
  } else {
    hipLaunchKernelGGL(pcg_kernel<float>, dim3(grid), dim3(kThreads), lds_bytes, stream, one, systems, items,
                       rounds, stride);
  }
  return hipGetLastError();
}

}  // namespace gbdpcg
