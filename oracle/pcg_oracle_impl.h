/*
 * Body of the CPU oracle, instantiated twice by pcg_oracle.c with
 *   REAL = double, SFX = f64   and   REAL = float, SFX = f32.
 * All arithmetic of an instance happens in REAL (the FP32 instance is the
 * "same precision" reference for iteration counts of the FP32 GPU path).
 * The exit comparisons are evaluated in double, as in the GPU kernel.
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* y_k = S_{k,k-1} x_{k-1} + S_{k,k} x_k + S_{k,k+1} x_{k+1} for every knot k.
 * Follows bdmv, include/utils.cuh:46-85: one output row r at a time, a sequential
 * sum over the columns of the n x 3n column-major block row (element (r,c) at
 * c*n + r); the first knot skips the (0,-1) block and the last knot the (N-1,N)
 * block (include/utils.cuh:50-65 and :66-74), i.e. the padding is never read.
 * The halo [x_{k-1}; x_k; x_{k+1}] is what loadbdVec gathers (include/utils.cuh:9-40). */
void FN(oracle_bdmv)(int n, int N, const REAL *A, const REAL *x, REAL *y) {
  const long bs = 3L * n * n;
  for (int k = 0; k < N; ++k) {
    const REAL *row = A + (long)k * bs;
    for (int r = 0; r < n; ++r) {
      REAL val = (REAL)0;
      const int c0 = k == 0 ? n : 0;           /* skip block (k, k-1) at the first knot */
      const int c1 = k == N - 1 ? 2 * n : 3 * n; /* skip block (k, k+1) at the last knot  */
      for (int c = c0; c < c1; ++c) {
        const long xi = (long)(k - 1) * n + c; /* column c of block row k is x entry (k-1)*n + c */
        val += row[(long)c * n + r] * x[xi];
      }
      y[(long)k * n + r] = val;
    }
  }
}

/* Inner product as a per-knot partial (glass::dot over n values, include/pcg.cuh:144)
 * followed by a sum of the N partials in knot order (glass::reduce, :147).  GLASS
 * is absent from the reference tree, so the order inside each is a choice: both
 * are sequential here. */
static REAL FN(dotk)(int n, int N, const REAL *a, const REAL *b) {
  REAL tot = (REAL)0;
  for (int k = 0; k < N; ++k) {
    REAL part = (REAL)0;
    for (int i = 0; i < n; ++i) part += a[(long)k * n + i] * b[(long)k * n + i];
    tot += part;
  }
  return tot;
}

static int FN(exit_met)(int mode, double tol, REAL eta, REAL rr, REAL gg) {
  if (mode == ORACLE_EXIT_REL) return (double)rr <= tol * tol * (double)gg;
  return fabs((double)eta) < tol;
}

/* The reference PCG iteration, include/pcg.cuh:118-208:
 *   r = gamma - S lambda; z = Pinv r; p = z; eta = r'z          (:118-149)
 *   for iter in [0, max_iter):                                   (:154)
 *     v = S p; alpha = eta / p'v                                 (:157-169)
 *     lambda += alpha p; r -= alpha v                            (:172-176)
 *     z = Pinv r; eta' = r'z                                     (:181-193)
 *     if |eta'| < tol: iter++, exit                              (:195)
 *     beta = eta'/eta; eta = eta'; p = z + beta p                (:199-206)
 *   iters = iter (max_iter when the test never passes)           (:212)
 * Extensions that do not change a run where they do not trigger: the
 * relative-residual exit mode, a zero-eta guard at the start (the reference
 * divides 0/0 there, :169), and breakdown/non-finite stops.
 * eta_hist (optional, max_iter+1 entries) receives eta_0, eta_1, ...
 * Returns the iteration count; *status gets an ORACLE_* status. */
int FN(oracle_pcg)(int n, int N, const REAL *S, const REAL *Pinv, const REAL *gamma, REAL *lambda, double tol,
                   int max_iter, int mode, double *eta_hist, int *status, double *eta_out, double *rr_out) {
  const long nn = (long)n * N;
  REAL *r = (REAL *)calloc(nn, sizeof(REAL)), *z = (REAL *)calloc(nn, sizeof(REAL));
  REAL *p = (REAL *)calloc(nn, sizeof(REAL)), *v = (REAL *)calloc(nn, sizeof(REAL));
  int st = ORACLE_MAX_ITER, iter = 0;
  FN(oracle_bdmv)(n, N, S, lambda, v);
  for (long i = 0; i < nn; ++i) r[i] = gamma[i] - v[i];
  const REAL gg = FN(dotk)(n, N, gamma, gamma);
  FN(oracle_bdmv)(n, N, Pinv, r, z);
  for (long i = 0; i < nn; ++i) p[i] = z[i];
  REAL eta = FN(dotk)(n, N, r, z);
  REAL rr = FN(dotk)(n, N, r, r);
  if (eta_hist) eta_hist[0] = (double)eta;
  double eta_last = (double)eta;
  if (!isfinite((double)eta) || !isfinite((double)rr)) {
    st = ORACLE_NONFINITE;
    goto done;
  }
  if (eta == (REAL)0) {
    st = rr == (REAL)0 ? ORACLE_CONVERGED : ORACLE_BREAKDOWN;
    goto done;
  }
  for (iter = 0; iter < max_iter;) {
    FN(oracle_bdmv)(n, N, S, p, v);
    const REAL pv = FN(dotk)(n, N, p, v);
    if (pv == (REAL)0 || !isfinite((double)pv)) {
      st = pv == (REAL)0 ? ORACLE_BREAKDOWN : ORACLE_NONFINITE;
      break;
    }
    const REAL alpha = eta / pv;
    for (long i = 0; i < nn; ++i) {
      lambda[i] += alpha * p[i];
      r[i] -= alpha * v[i];
    }
    FN(oracle_bdmv)(n, N, Pinv, r, z);
    const REAL eta_new = FN(dotk)(n, N, r, z);
    rr = FN(dotk)(n, N, r, r);
    ++iter;
    eta_last = (double)eta_new;
    if (eta_hist) eta_hist[iter] = (double)eta_new;
    if (FN(exit_met)(mode, tol, eta_new, rr, gg)) {
      st = ORACLE_CONVERGED;
      break;
    }
    if (!isfinite((double)eta_new)) {
      st = ORACLE_NONFINITE;
      break;
    }
    const REAL beta = eta_new / eta;
    eta = eta_new;
    for (long i = 0; i < nn; ++i) p[i] = z[i] + beta * p[i];
  }
done:
  if (status) *status = st;
  if (eta_out) *eta_out = eta_last;
  if (rr_out) *rr_out = (double)rr;
  free(r);
  free(z);
  free(p);
  free(v);
  return iter;
}

/* The same method in the Chronopoulos-Gear single-reduction form (A. T.
 * Chronopoulos and C. W. Gear, J. Comput. Appl. Math. 25 (1989) 153-168):
 *   r = b - A x; u = M r; w = A u; gamma = (r,u); delta = (w,u); alpha = gamma/delta
 *   p = u; s = w
 *   loop: x += alpha p; r -= alpha s; u = M r; w = A u;
 *         gamma' = (r,u); delta' = (w,u); [exit on gamma' as in the reference]
 *         beta = gamma'/gamma; alpha = gamma' / (delta' - beta gamma'/alpha)
 *         p = u + beta p; s = w + beta s
 * Identical iterates to oracle_pcg in exact arithmetic; counted the same way. */
int FN(oracle_pcg_fused)(int n, int N, const REAL *S, const REAL *Pinv, const REAL *gamma, REAL *lambda, double tol,
                         int max_iter, int mode, double *eta_hist, int *status, double *eta_out, double *rr_out) {
  const long nn = (long)n * N;
  REAL *r = (REAL *)calloc(nn, sizeof(REAL)), *u = (REAL *)calloc(nn, sizeof(REAL));
  REAL *w = (REAL *)calloc(nn, sizeof(REAL)), *p = (REAL *)calloc(nn, sizeof(REAL));
  REAL *s = (REAL *)calloc(nn, sizeof(REAL));
  int st = ORACLE_MAX_ITER, iter = 0;
  FN(oracle_bdmv)(n, N, S, lambda, w);
  for (long i = 0; i < nn; ++i) r[i] = gamma[i] - w[i];
  const REAL gg = FN(dotk)(n, N, gamma, gamma);
  FN(oracle_bdmv)(n, N, Pinv, r, u);
  FN(oracle_bdmv)(n, N, S, u, w);
  REAL eta = FN(dotk)(n, N, r, u);
  REAL delta = FN(dotk)(n, N, w, u);
  REAL rr = FN(dotk)(n, N, r, r);
  REAL alpha = eta / delta;
  for (long i = 0; i < nn; ++i) {
    p[i] = u[i];
    s[i] = w[i];
  }
  if (eta_hist) eta_hist[0] = (double)eta;
  double eta_last = (double)eta;
  if (!isfinite((double)eta) || !isfinite((double)rr)) {
    st = ORACLE_NONFINITE;
    goto done;
  }
  if (eta == (REAL)0) {
    st = rr == (REAL)0 ? ORACLE_CONVERGED : ORACLE_BREAKDOWN;
    goto done;
  }
  if (delta == (REAL)0 || !isfinite((double)alpha)) {
    st = ORACLE_BREAKDOWN;
    goto done;
  }
  for (iter = 0; iter < max_iter;) {
    for (long i = 0; i < nn; ++i) {
      lambda[i] += alpha * p[i];
      r[i] -= alpha * s[i];
    }
    FN(oracle_bdmv)(n, N, Pinv, r, u);
    FN(oracle_bdmv)(n, N, S, u, w);
    const REAL eta_new = FN(dotk)(n, N, r, u);
    const REAL delta_new = FN(dotk)(n, N, w, u);
    rr = FN(dotk)(n, N, r, r);
    ++iter;
    eta_last = (double)eta_new;
    if (eta_hist) eta_hist[iter] = (double)eta_new;
    if (FN(exit_met)(mode, tol, eta_new, rr, gg)) {
      st = ORACLE_CONVERGED;
      break;
    }
    if (!isfinite((double)eta_new)) {
      st = ORACLE_NONFINITE;
      break;
    }
    const REAL beta = eta_new / eta;
    const REAL den = delta_new - beta * eta_new / alpha;
    if (den == (REAL)0 || !isfinite((double)den)) {
      st = ORACLE_BREAKDOWN;
      break;
    }
    alpha = eta_new / den;
    eta = eta_new;
    for (long i = 0; i < nn; ++i) {
      p[i] = u[i] + beta * p[i];
      s[i] = w[i] + beta * s[i];
    }
  }
done:
  if (status) *status = st;
  if (eta_out) *eta_out = eta_last;
  if (rr_out) *rr_out = (double)rr;
  free(r);
  free(u);
  free(w);
  free(p);
  free(s);
  return iter;
}

#undef FN
#undef CAT
#undef CAT2
