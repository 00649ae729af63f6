// Host-side problem construction: the seeded LQR Schur-complement generator,
// the block-Jacobi / symmetric-stair preconditioner on the host, and CSR -> cbtd.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "internal.h"

namespace gbdpcg {

void set_error(const char *fmt, ...);  // api.cpp

namespace {

// splitmix64 (Steele, Lea, Flood 2014) seeds xoshiro256** (Blackman & Vigna 2018).
struct Rng {
  uint64_t s[4];
  bool has_spare = false;
  double spare = 0.0;

  static uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &w : s) w = splitmix64(x);
  }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // uniform in [0, 1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // standard normal, Box-Muller (1958) polar-free form, both outputs used
  double normal() {
    if (has_spare) {
      has_spare = false;
      return spare;
    }
    double u1;
    do {
      u1 = uniform();
    } while (u1 <= 0.0);
    const double u2 = uniform();
    const double rad = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586476925286766559 * u2;
    spare = rad * sin(th);
    has_spare = true;
    return rad * cos(th);
  }
};

void put(int dtype, void *dst, size_t i, double v) {
  if (dtype == GBDPCG_F64)
    static_cast<double *>(dst)[i] = v;
  else
    static_cast<float *>(dst)[i] = (float)v;
}

double get(int dtype, const void *src, size_t i) {
  return dtype == GBDPCG_F64 ? static_cast<const double *>(src)[i] : (double)static_cast<const float *>(src)[i];
}

// In-place Gauss-Jordan inverse with partial pivoting, row-major n x n.
bool invert(std::vector<double> &a, int n) {
  std::vector<double> aug((size_t)n * 2 * n, 0.0);
  for (int r = 0; r < n; ++r) {
    for (int c = 0; c < n; ++c) aug[(size_t)r * 2 * n + c] = a[(size_t)r * n + c];
    aug[(size_t)r * 2 * n + n + r] = 1.0;
  }
  for (int col = 0; col < n; ++col) {
    int best = col;
    double bv = fabs(aug[(size_t)col * 2 * n + col]);
    for (int r = col + 1; r < n; ++r) {
      const double v = fabs(aug[(size_t)r * 2 * n + col]);
      if (v > bv) {
        bv = v;
        best = r;
      }
    }
    if (bv == 0.0) return false;
    if (best != col)
      for (int j = 0; j < 2 * n; ++j) std::swap(aug[(size_t)col * 2 * n + j], aug[(size_t)best * 2 * n + j]);
    const double ip = 1.0 / aug[(size_t)col * 2 * n + col];
    for (int j = 0; j < 2 * n; ++j) aug[(size_t)col * 2 * n + j] *= ip;
    for (int r = 0; r < n; ++r) {
      if (r == col) continue;
      const double f = aug[(size_t)r * 2 * n + col];
      if (f == 0.0) continue;
      for (int j = 0; j < 2 * n; ++j) aug[(size_t)r * 2 * n + j] -= f * aug[(size_t)col * 2 * n + j];
    }
  }
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) a[(size_t)r * n + c] = aug[(size_t)r * 2 * n + n + c];
  return true;
}

}  // namespace

}  // namespace gbdpcg

using namespace gbdpcg;

extern "C" int gbdpcg_randn(int dtype, uint64_t seed, uint64_t count, double scale, void *h_out) {
  if ((dtype != GBDPCG_F32 && dtype != GBDPCG_F64) || (count && !h_out)) {
    set_error("gbdpcg_randn: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  Rng rng(seed);
  for (uint64_t i = 0; i < count; ++i) put(dtype, h_out, i, scale * rng.normal());
  return GBDPCG_OK;
}

extern "C" int gbdpcg_counter_randn(int dtype, uint64_t seed, uint64_t stream, uint64_t count, double scale,
                                    void *h_out) {
  if ((dtype != GBDPCG_F32 && dtype != GBDPCG_F64) || (count && !h_out)) {
    set_error("gbdpcg_counter_randn: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  const uint64_t key = counter_key(seed, stream);
  for (uint64_t i = 0; i < count; ++i) put(dtype, h_out, i, scale * counter_normal(key, i));
  return GBDPCG_OK;
}

extern "C" int gbdpcg_lqr_schur_host(int dtype, uint32_t n, uint32_t m, uint32_t N, const double *A,
                                     const double *B, const double *Qd, const double *Rd, void *h_S) {
  if ((dtype != GBDPCG_F32 && dtype != GBDPCG_F64) || n == 0 || N == 0 || !A || (m && !B) || !Qd ||
      (m && !Rd) || !h_S) {
    set_error("gbdpcg_lqr_schur_host: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  const size_t bs = (size_t)3 * n * n;
  for (uint32_t k = 0; k < N; ++k) {
    for (uint32_t e = 0; e < 3 * n * n; ++e) {
      const uint32_t col = e / n, r = e % n, seg = col / n, c = col % n;
      double val = 0.0;
      if (seg == 0) {
        if (k > 0) val = -A[(size_t)(k - 1) * n * n + r * n + c] / Qd[(size_t)(k - 1) * n + c];
      } else if (seg == 1) {
        if (k > 0) {
          const double *Ak = A + (size_t)(k - 1) * n * n;
          const double *qk = Qd + (size_t)(k - 1) * n;
          for (uint32_t i = 0; i < n; ++i) val += Ak[r * n + i] * Ak[c * n + i] / qk[i];
          if (m) {
            const double *Bk = B + (size_t)(k - 1) * n * m;
            const double *rk = Rd + (size_t)(k - 1) * m;
            for (uint32_t j = 0; j < m; ++j) val += Bk[r * m + j] * Bk[c * m + j] / rk[j];
          }
        }
        if (r == c) val += 1.0 / Qd[(size_t)k * n + r];
      } else {
        if (k + 1 < N) val = -A[(size_t)k * n * n + c * n + r] / Qd[(size_t)k * n + r];
      }
      put(dtype, h_S, (size_t)k * bs + e, val);
    }
  }
  return GBDPCG_OK;
}

extern "C" int gbdpcg_lqr_generate(int dtype, uint32_t n, uint32_t m, uint32_t N, uint64_t seed, double dt,
                                   void *h_S, void *h_gamma, double *A, double *B, double *Qd, double *Rd) {
  if ((dtype != GBDPCG_F32 && dtype != GBDPCG_F64) || n == 0 || N == 0) {
    set_error("gbdpcg_lqr_generate: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  std::vector<double> a((size_t)N * n * n), b((size_t)N * n * m), q((size_t)N * n), r((size_t)N * m);
  Rng rng(seed);
  for (uint32_t k = 0; k < N; ++k) {
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t j = 0; j < n; ++j)
        a[(size_t)k * n * n + i * n + j] = (i == j ? 1.0 : 0.0) + dt * (0.1 * rng.normal());
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t j = 0; j < m; ++j) b[(size_t)k * n * m + i * m + j] = dt * rng.normal();
    for (uint32_t i = 0; i < n; ++i) q[(size_t)k * n + i] = 0.1 + 9.9 * rng.uniform();
    for (uint32_t j = 0; j < m; ++j) r[(size_t)k * m + j] = 0.01 + 0.99 * rng.uniform();
  }
  if (h_gamma)
    for (size_t i = 0; i < (size_t)n * N; ++i) put(dtype, h_gamma, i, rng.normal());
  if (h_S) {
    const int rc = gbdpcg_lqr_schur_host(dtype, n, m, N, a.data(), m ? b.data() : nullptr, q.data(),
                                         m ? r.data() : nullptr, h_S);
    if (rc < 0) return rc;
  }
  if (A) memcpy(A, a.data(), a.size() * sizeof(double));
  if (B && m) memcpy(B, b.data(), b.size() * sizeof(double));
  if (Qd) memcpy(Qd, q.data(), q.size() * sizeof(double));
  if (Rd && m) memcpy(Rd, r.data(), r.size() * sizeof(double));
  return GBDPCG_OK;
}

extern "C" int gbdpcg_precond_host(int dtype, uint32_t n, uint32_t N, const void *h_S, void *h_Pinv, int kind) {
  if ((dtype != GBDPCG_F32 && dtype != GBDPCG_F64) || n == 0 || N == 0 || !h_S || !h_Pinv ||
      (kind != GBDPCG_PRECOND_JACOBI && kind != GBDPCG_PRECOND_SS)) {
    set_error("gbdpcg_precond_host: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  const size_t bs = (size_t)3 * n * n, nn = (size_t)n * n;
  // inverses of all diagonal blocks (row-major)
  std::vector<std::vector<double>> inv(N, std::vector<double>(nn));
  for (uint32_t k = 0; k < N; ++k) {
    for (uint32_t r = 0; r < n; ++r)
      for (uint32_t c = 0; c < n; ++c) inv[k][r * n + c] = get(dtype, h_S, k * bs + nn + (size_t)c * n + r);
    if (!invert(inv[k], (int)n)) {
      set_error("gbdpcg_precond_host: diagonal block %u is singular", k);
      return GBDPCG_ERR_INVALID;
    }
  }
  std::vector<double> phi(nn), t(nn), u(nn);
  auto mm = [&](const std::vector<double> &X, const std::vector<double> &Y, std::vector<double> &Z) {
    for (uint32_t r = 0; r < n; ++r)
      for (uint32_t c = 0; c < n; ++c) {
        double acc = 0.0;
        for (uint32_t q = 0; q < n; ++q) acc += X[r * n + q] * Y[q * n + c];
        Z[r * n + c] = acc;
      }
  };
  for (uint32_t k = 0; k < N; ++k) {
    for (size_t e = 0; e < bs; ++e) put(dtype, h_Pinv, k * bs + e, 0.0);
    for (uint32_t r = 0; r < n; ++r)
      for (uint32_t c = 0; c < n; ++c) put(dtype, h_Pinv, k * bs + nn + (size_t)c * n + r, inv[k][r * n + c]);
    if (kind != GBDPCG_PRECOND_SS) continue;
    if (k > 0) {  // (k,k-1) = -inv_k phi_k inv_{k-1}
      for (uint32_t r = 0; r < n; ++r)
        for (uint32_t c = 0; c < n; ++c) phi[r * n + c] = get(dtype, h_S, k * bs + (size_t)c * n + r);
      mm(inv[k], phi, t);
      mm(t, inv[k - 1], u);
      for (uint32_t r = 0; r < n; ++r)
        for (uint32_t c = 0; c < n; ++c) put(dtype, h_Pinv, k * bs + (size_t)c * n + r, -u[r * n + c]);
    }
    if (k + 1 < N) {  // (k,k+1) = (-inv_{k+1} phi_{k+1} inv_k)^T
      for (uint32_t r = 0; r < n; ++r)
        for (uint32_t c = 0; c < n; ++c) phi[r * n + c] = get(dtype, h_S, (k + 1) * bs + (size_t)c * n + r);
      mm(inv[k + 1], phi, t);
      mm(t, inv[k], u);
      for (uint32_t r = 0; r < n; ++r)
        for (uint32_t c = 0; c < n; ++c) put(dtype, h_Pinv, k * bs + 2 * nn + (size_t)c * n + r, -u[c * n + r]);
    }
  }
  return GBDPCG_OK;
}

extern "C" int gbdpcg_csr_to_cbtd(int dtype, const uint32_t *row_ptr, const uint32_t *col_ind, const void *val,
                                  uint32_t n, uint32_t N, void *cbtd_out) {
  if ((dtype != GBDPCG_F32 && dtype != GBDPCG_F64) || !row_ptr || !cbtd_out || n == 0 || N == 0) {
    set_error("gbdpcg_csr_to_cbtd: invalid argument");
    return GBDPCG_ERR_INVALID;
  }
  const size_t rows = (size_t)n * N, bs = (size_t)3 * n * n;
  const size_t nnz = row_ptr[rows];
  if (nnz && (!col_ind || !val)) {
    set_error("gbdpcg_csr_to_cbtd: null col_ind/val with nnz > 0");
    return GBDPCG_ERR_INVALID;
  }
  for (size_t i = 0; i < bs * N; ++i) put(dtype, cbtd_out, i, 0.0);
  for (size_t i = 0; i < rows; ++i) {
    if (row_ptr[i + 1] < row_ptr[i]) {
      set_error("gbdpcg_csr_to_cbtd: row_ptr not monotone at row %zu", i);
      return GBDPCG_ERR_INVALID;
    }
    const size_t k = i / n, r = i % n;
    for (uint32_t q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
      const size_t j = col_ind[q];
      if (j >= rows) {
        set_error("gbdpcg_csr_to_cbtd: column %zu out of range", j);
        return GBDPCG_ERR_INVALID;
      }
      const long seg = (long)(j / n) - (long)k + 1;
      const double v = get(dtype, val, q);
      if (seg < 0 || seg > 2) {
        if (v == 0.0) continue;  // explicit zeros outside the pattern are harmless
        set_error("gbdpcg_csr_to_cbtd: entry (%zu,%zu) outside the block-tridiagonal pattern", i, j);
        return GBDPCG_ERR_UNSUPPORTED;
      }
      const size_t c = j % n;
      const size_t idx = k * bs + ((size_t)seg * n + c) * n + r;
      put(dtype, cbtd_out, idx, get(dtype, cbtd_out, idx) + v);  // duplicates are summed
    }
  }
  return GBDPCG_OK;
}
