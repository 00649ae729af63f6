/*
 * CPU oracle for the block-tridiagonal PCG path — TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker (or the timed CPU baseline); the product
 * never calls it.  It restates, in plain C, the algorithm of the reference
 * (Sdevaud/GBD-PCG): the PCG kernel include/pcg.cuh:103-217, the block-row matvec
 * include/utils.cuh:46-85 with the halo of include/utils.cuh:9-40, and the
 * iteration-count semantics of include/interface.cuh:94-146.  The reference itself
 * cannot be built here (CUDA only, and its GLASS dependency — A2R-Lab/GLASS, an
 * empty submodule with an unrecoverable pinned commit — is absent), so the oracle
 * is pinned by the checks in tests/test_oracle.py (closed forms, a dense solve,
 * SciPy's independent PCG, invariants and the reference example's data).
 *
 * Also here: a block-tridiagonal direct solve (block LU / "block Thomas" with
 * partial-pivoting dense LU on each diagonal block) used as the FP64 solution
 * oracle when a dense solve is too large.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_EXIT_ABS 0
#define ORACLE_EXIT_REL 1
#define ORACLE_CONVERGED 0
#define ORACLE_MAX_ITER 1
#define ORACLE_BREAKDOWN 2
#define ORACLE_NONFINITE 3

#define REAL double
#define SFX f64
#include "pcg_oracle_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX f32
#include "pcg_oracle_impl.h"
#undef REAL
#undef SFX

/* dense LU with partial pivoting of an m x m row-major matrix (in place);
 * returns 0 on success, -1 on an exactly zero pivot. */
static int lu_factor(int m, double *a, int *piv) {
  for (int c = 0; c < m; ++c) {
    int best = c;
    double bv = fabs(a[(long)c * m + c]);
    for (int r = c + 1; r < m; ++r)
      if (fabs(a[(long)r * m + c]) > bv) {
        bv = fabs(a[(long)r * m + c]);
        best = r;
      }
    piv[c] = best;
    if (bv == 0.0) return -1;
    if (best != c)
      for (int j = 0; j < m; ++j) {
        const double t = a[(long)c * m + j];
        a[(long)c * m + j] = a[(long)best * m + j];
        a[(long)best * m + j] = t;
      }
    for (int r = c + 1; r < m; ++r) {
      const double f = a[(long)r * m + c] / a[(long)c * m + c];
      a[(long)r * m + c] = f;
      for (int j = c + 1; j < m; ++j) a[(long)r * m + j] -= f * a[(long)c * m + j];
    }
  }
  return 0;
}

/* solves LU x = b for `cols` right-hand sides stored row-major in b (m x cols).
 * lu_factor swaps whole rows (the stored multipliers move with them), so all the
 * interchanges are applied to b first, then plain forward substitution. */
static void lu_solve(int m, const double *lu, const int *piv, double *b, int cols) {
  for (int c = 0; c < m; ++c)
    if (piv[c] != c)
      for (int j = 0; j < cols; ++j) {
        const double t = b[(long)c * cols + j];
        b[(long)c * cols + j] = b[(long)piv[c] * cols + j];
        b[(long)piv[c] * cols + j] = t;
      }
  for (int c = 0; c < m; ++c) {
    for (int r = c + 1; r < m; ++r)
      for (int j = 0; j < cols; ++j) b[(long)r * cols + j] -= lu[(long)r * m + c] * b[(long)c * cols + j];
  }
  for (int c = m - 1; c >= 0; --c)
    for (int j = 0; j < cols; ++j) {
      double v = b[(long)c * cols + j];
      for (int q = c + 1; q < m; ++q) v -= lu[(long)c * m + q] * b[(long)q * cols + j];
      b[(long)c * cols + j] = v / lu[(long)c * m + c];
    }
}

/* Block-tridiagonal solve S x = b (cbtd S, FP64) by block LU without pivoting
 * across block rows:  D_0 = S_00;  D_k = S_kk - S_{k,k-1} D_{k-1}^-1 S_{k-1,k};
 * forward y_k = b_k - S_{k,k-1} D_{k-1}^-1 y_{k-1}; backward
 * x_k = D_k^-1 (y_k - S_{k,k+1} x_{k+1}).  Stable for definite S (either sign).
 * Returns 0, or -1 if a D_k is singular. */
int oracle_block_solve_f64(int n, int N, const double *S, const double *b, double *x) {
  const long bs = 3L * n * n, nn = (long)n * n;
  double *D = (double *)malloc(sizeof(double) * nn * N);  /* LU of D_k */
  int *piv = (int *)malloc(sizeof(int) * n * N);
  double *y = (double *)malloc(sizeof(double) * n * N);
  double *L = (double *)malloc(sizeof(double) * nn);      /* scratch: D_{k-1}^-1 S_{k-1,k} */
  int rc = 0;
#define BLK(k, seg, r, c) S[(long)(k) * bs + (long)(seg) * nn + (long)(c) * n + (r)]
  for (int k = 0; k < N && rc == 0; ++k) {
    double *Dk = D + (long)k * nn;
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) Dk[(long)r * n + c] = BLK(k, 1, r, c);
    for (int r = 0; r < n; ++r) y[(long)k * n + r] = b[(long)k * n + r];
    if (k > 0) {
      /* L = D_{k-1}^-1 S_{k-1,k}  (n x n), then D_k -= S_{k,k-1} L */
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) L[(long)r * n + c] = BLK(k - 1, 2, r, c);
      lu_solve(n, D + (long)(k - 1) * nn, piv + (long)(k - 1) * n, L, n);
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
          double acc = 0.0;
          for (int q = 0; q < n; ++q) acc += BLK(k, 0, r, q) * L[(long)q * n + c];
          Dk[(long)r * n + c] -= acc;
        }
      /* y_k -= S_{k,k-1} D_{k-1}^-1 y_{k-1} */
      double *t = (double *)malloc(sizeof(double) * n);
      for (int r = 0; r < n; ++r) t[r] = y[(long)(k - 1) * n + r];
      lu_solve(n, D + (long)(k - 1) * nn, piv + (long)(k - 1) * n, t, 1);
      for (int r = 0; r < n; ++r) {
        double acc = 0.0;
        for (int q = 0; q < n; ++q) acc += BLK(k, 0, r, q) * t[q];
        y[(long)k * n + r] -= acc;
      }
      free(t);
    }
    if (lu_factor(n, Dk, piv + (long)k * n) != 0) rc = -1;
  }
  if (rc == 0) {
    for (int k = N - 1; k >= 0; --k) {
      double *t = x + (long)k * n;
      for (int r = 0; r < n; ++r) {
        double v = y[(long)k * n + r];
        if (k + 1 < N)
          for (int q = 0; q < n; ++q) v -= BLK(k, 2, r, q) * x[(long)(k + 1) * n + q];
        t[r] = v;
      }
      lu_solve(n, D + (long)k * nn, piv + (long)k * n, t, 1);
    }
  }
#undef BLK
  free(D);
  free(piv);
  free(y);
  free(L);
  return rc;
}
