// Auxiliary device kernels: fills, preconditioner formation from S, LQR Schur
// complement formation, and the receding-horizon warm start.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"
#include "internal.h"

namespace gbdpcg {

__global__ void fill_u64_kernel(uint64_t *p, size_t count, uint64_t value) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    p[i] = value;
}

__global__ void fill_i32_kernel(int *p, size_t count, int value) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    p[i] = value;
}

hipError_t fill_u64(uint64_t *p, size_t count, uint64_t value, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  size_t blocks = (count + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, count, value);
  return hipGetLastError();
}

hipError_t fill_i32(int *p, size_t count, int value, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  size_t blocks = (count + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, count, value);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// preconditioner: block-Jacobi or symmetric stair, one workgroup per knot
// ---------------------------------------------------------------------------

constexpr int kAuxThreads = 256;

// In-place Gauss-Jordan inversion with partial pivoting of the n x n row-major
// matrix held in the left half of aug (n x 2n, right half set to I here).
// Returns false (uniformly) if a pivot is exactly zero.
template <typename T>
__device__ bool gj_invert(T *aug, T *fac, int *piv, int n) {
  const int n2 = 2 * n;
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) {
    const int r = i / n, c = i - r * n;
    aug[r * n2 + n + c] = r == c ? (T)1 : (T)0;
  }
  __syncthreads();
  for (int col = 0; col < n; ++col) {
    if (threadIdx.x == 0) {
      int best = col;
      T bv = __builtin_fabs((double)aug[col * n2 + col]);
      for (int r = col + 1; r < n; ++r) {
        const T v = __builtin_fabs((double)aug[r * n2 + col]);
        if (v > bv) {
          bv = v;
          best = r;
        }
      }
      piv[0] = best;
      piv[1] = bv == (T)0 ? 1 : 0;
    }
    __syncthreads();
    if (piv[1]) return false;
    const int pr = piv[0];
    if (pr != col) {
      for (int j = threadIdx.x; j < n2; j += blockDim.x) {
        const T t = aug[col * n2 + j];
        aug[col * n2 + j] = aug[pr * n2 + j];
        aug[pr * n2 + j] = t;
      }
    }
    __syncthreads();
    const T inv_p = (T)1 / aug[col * n2 + col];
    __syncthreads();
    for (int j = threadIdx.x; j < n2; j += blockDim.x) aug[col * n2 + j] *= inv_p;
    for (int r = threadIdx.x; r < n; r += blockDim.x) fac[r] = r == col ? (T)0 : aug[r * n2 + col];
    __syncthreads();
    for (int i = threadIdx.x; i < n * n2; i += blockDim.x) {
      const int r = i / n2, j = i - r * n2;
      if (r != col) aug[i] -= fac[r] * aug[col * n2 + j];
    }
    __syncthreads();
  }
  return true;
}

// dst (n x n row-major) = theta_kk^-1, theta read column-major from S block (k,k).
template <typename T>
__device__ bool invert_diag(const T *S, int k, int n, T *aug, T *fac, int *piv, T *dst) {
  const size_t bs = (size_t)3 * n * n;
  const int n2 = 2 * n;
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) {
    const int r = i / n, c = i - r * n;
    aug[r * n2 + c] = S[(size_t)k * bs + (size_t)n * n + (size_t)c * n + r];
  }
  __syncthreads();
  if (!gj_invert(aug, fac, piv, n)) return false;
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) {
    const int r = i / n, c = i - r * n;
    dst[i] = aug[r * n2 + n + c];
  }
  __syncthreads();
  return true;
}

// C = A * B, all n x n row-major in LDS.
template <typename T>
__device__ void mm(const T *A, const T *B, T *C, int n) {
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) {
    const int r = i / n, c = i - r * n;
    T acc = 0;
    for (int q = 0; q < n; ++q) acc += A[r * n + q] * B[q * n + c];
    C[i] = acc;
  }
  __syncthreads();
}

// Pinv row k.  LDS: aug 2n^2 | inv_prev n^2 | inv_k n^2 | inv_next n^2 | t1 n^2 | t2 n^2 | fac n.
template <typename T>
__global__ void __launch_bounds__(kAuxThreads) precond_kernel(const T *__restrict__ S, T *__restrict__ P, int n,
                                                             int N, int kind, int *err) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  T *aug = reinterpret_cast<T *>(lds_raw);
  T *inv_prev = aug + 2 * n * n;
  T *inv_k = inv_prev + n * n;
  T *inv_next = inv_k + n * n;
  T *t1 = inv_next + n * n;
  T *t2 = t1 + n * n;
  T *fac = t2 + n * n;
  __shared__ int piv[2];
  const int k = blockIdx.x;
  const size_t bs = (size_t)3 * n * n;
  T *row = P + (size_t)k * bs;
  bool ok = invert_diag(S, k, n, aug, fac, piv, inv_k);
  const bool ss = kind == GBDPCG_PRECOND_SS;
  if (ok && ss && k > 0) ok = invert_diag(S, k - 1, n, aug, fac, piv, inv_prev);
  if (ok && ss && k < N - 1) ok = invert_diag(S, k + 1, n, aug, fac, piv, inv_next);
  if (!ok) {
    if (threadIdx.x == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    for (int i = threadIdx.x; i < (int)bs; i += blockDim.x) row[i] = (T)0;
    return;
  }
  // diagonal block (column-major)
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) {
    const int c = i / n, r = i - c * n;
    row[(size_t)n * n + i] = inv_k[r * n + c];
  }
  // (k,k-1) = -inv_k phi_k inv_prev, phi_k = S(k,k-1)
  if (ss && k > 0) {
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) {
      const int r = i / n, c = i - r * n;
      t1[i] = S[(size_t)k * bs + (size_t)c * n + r];
    }
    __syncthreads();
    mm(inv_k, t1, t2, n);
    mm(t2, inv_prev, t1, n);
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) {
      const int c = i / n, r = i - c * n;
      row[i] = -t1[r * n + c];
    }
  } else {
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) row[i] = (T)0;
  }
  __syncthreads();
  // (k,k+1) = transpose of (k+1,k) = (-inv_next phi_{k+1} inv_k)^T
  if (ss && k < N - 1) {
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) {
      const int r = i / n, c = i - r * n;
      t1[i] = S[(size_t)(k + 1) * bs + (size_t)c * n + r];
    }
    __syncthreads();
    mm(inv_next, t1, t2, n);
    mm(t2, inv_k, t1, n);
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) {
      const int c = i / n, r = i - c * n;
      row[2 * (size_t)n * n + i] = -t1[c * n + r];
    }
  } else {
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) row[2 * (size_t)n * n + i] = (T)0;
  }
}

// ---------------------------------------------------------------------------
// preconditioner, n <= 32: two launches
//   inv_diag_kernel   one wave per knot inverts theta_k in registers and writes
//                     it to the diagonal block of Pinv row k;
//   stair_kernel      one workgroup per knot forms the symmetric-stair
//                     off-diagonal blocks from those inverses (or zeros them for
//                     block-Jacobi).
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T lane_max(T v) {  // over the wave; v >= 0 or -1 (excluded)
  v = __builtin_fmax(v, dpp<kDppRowShr1>(v));
  v = __builtin_fmax(v, dpp<kDppRowShr2>(v));
  v = __builtin_fmax(v, dpp<kDppRowShr4>(v));
  v = __builtin_fmax(v, dpp<kDppRowShr8>(v));
  v = __builtin_fmax(v, dpp<kDppRowBcast15, 0xa>(v));
  v = __builtin_fmax(v, dpp<kDppRowBcast31, 0xc>(v));
  return read_lane63(v);
}

template <typename T>
__device__ __forceinline__ T lane_read(T v, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
  }
}

// Gauss-Jordan with partial pivoting on [theta | I], lane r holding row r.  Rows
// are never swapped: the lane chosen as pivot for column c is marked used, so at
// the end its left half is the unit vector e_c and its right half is row c of the
// inverse.  Ties in the pivot search go to the lowest row, as in a sequential
// search for the first maximum.
template <typename T, int NM>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) inv_diag_kernel(const T *__restrict__ S, T *__restrict__ P, int n, int N,
                                                     int *err) {
  const int k = blockIdx.x;
  const int lane = threadIdx.x;
  const size_t bs = (size_t)3 * n * n;
  const bool is_row = lane < n;
  T al[NM], ar[NM];  // [theta | I] row `lane`: left and right halves (two arrays stay in registers)
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    al[j] = (is_row && j < n) ? S[(size_t)k * bs + (size_t)n * n + (size_t)j * n + lane] : (T)0;
    ar[j] = (is_row && j == lane) ? (T)1 : (T)0;
  }
  bool used = !is_row;
  int mycol = 0;
  bool ok = true;
#pragma unroll
  for (int col = 0; col < NM; ++col) {
    if (col < n && ok) {
      const T cand = used ? (T)-1 : __builtin_fabs(al[col]);
      const T m = lane_max(cand);
      if (!(m > (T)0)) {
        ok = false;  // exactly zero (or NaN) pivot: singular block
      } else {
        const unsigned long long hit = __ballot(!used && cand == m);
        const int pr = __builtin_amdgcn_readfirstlane(__ffsll((long long)hit) - 1);
        const T ip = (T)1 / lane_read(al[col], pr);
        if (lane == pr) {
#pragma unroll
          for (int j = 0; j < NM; ++j) {
            al[j] *= ip;
            ar[j] *= ip;
          }
          used = true;
          mycol = col;
        }
        const T f = lane == pr ? (T)0 : al[col];
#pragma unroll
        for (int j = 0; j < NM; ++j) {
          if (j < n) {
            al[j] = fmaT(-f, lane_read(al[j], pr), al[j]);
            ar[j] = fmaT(-f, lane_read(ar[j], pr), ar[j]);
          }
        }
      }
    }
  }
  T *row = P + (size_t)k * bs;
  if (!ok) {
    if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    for (int i = lane; i < (int)bs; i += 64) row[i] = (T)0;
    return;
  }
  // diagonal block, column-major: element (mycol, j) of the inverse at j*n + mycol
  if (is_row) {
#pragma unroll
    for (int j = 0; j < NM; ++j)
      if (j < n) row[(size_t)n * n + (size_t)j * n + mycol] = ar[j];
  }
}

// LDS: inv_k | inv_nb | phi | t  (n x n row-major each).
template <typename T>
__global__ void __launch_bounds__(kAuxThreads) stair_kernel(const T *__restrict__ S, T *__restrict__ P, int n, int N,
                                                           int kind) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  T *inv_k = reinterpret_cast<T *>(lds_raw);
  T *inv_nb = inv_k + n * n;
  T *phi = inv_nb + n * n;
  T *t = phi + n * n;
  const int k = blockIdx.x;
  const size_t bs = (size_t)3 * n * n, nn = (size_t)n * n;
  T *row = P + (size_t)k * bs;
  const bool ss = kind == GBDPCG_PRECOND_SS;
  if (!ss) {
    for (int i = threadIdx.x; i < (int)nn; i += blockDim.x) {
      row[i] = (T)0;
      row[2 * nn + i] = (T)0;
    }
    return;
  }
  // inverses are stored column-major in the diagonal blocks
  for (int i = threadIdx.x; i < (int)nn; i += blockDim.x) {
    const int c = i / n, r = i - c * n;
    inv_k[r * n + c] = P[(size_t)k * bs + nn + i];
  }
  for (int side = 0; side < 2; ++side) {
    // side 0: (k,k-1) = -inv_k phi_k inv_{k-1};  side 1: (k,k+1) = (-inv_{k+1} phi_{k+1} inv_k)^T
    const int nb = side == 0 ? k - 1 : k + 1;
    T *dst = row + (side == 0 ? 0 : 2 * nn);
    if (nb < 0 || nb >= N) {
      for (int i = threadIdx.x; i < (int)nn; i += blockDim.x) dst[i] = (T)0;
      continue;
    }
    const int kp = side == 0 ? k : k + 1;  // phi_kp = S(kp, kp-1)
    __syncthreads();
    for (int i = threadIdx.x; i < (int)nn; i += blockDim.x) {
      const int c = i / n, r = i - c * n;
      inv_nb[r * n + c] = P[(size_t)nb * bs + nn + i];
      phi[r * n + c] = S[(size_t)kp * bs + i];
    }
    __syncthreads();
    const T *left = side == 0 ? inv_k : inv_nb;   // inv of kp
    const T *right = side == 0 ? inv_nb : inv_k;  // inv of kp-1
    for (int i = threadIdx.x; i < (int)nn; i += blockDim.x) {  // t = phi * right
      const int r = i / n, c = i - r * n;
      T acc = 0;
      for (int q = 0; q < n; ++q) acc += phi[r * n + q] * right[q * n + c];
      t[i] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < (int)nn; i += blockDim.x) {  // -(left * t), column-major (or transposed)
      const int r = i / n, c = i - r * n;
      T acc = 0;
      for (int q = 0; q < n; ++q) acc += left[r * n + q] * t[q * n + c];
      if (side == 0)
        dst[(size_t)c * n + r] = -acc;
      else
        dst[(size_t)r * n + c] = -acc;  // transpose: element (c, r) of block (k,k+1), column-major
    }
  }
}

template <typename T>
static hipError_t precond_small(uint32_t n, uint32_t N, const T *S, T *P, int kind, int *err, hipStream_t stream) {
  if (n <= 4)
    hipLaunchKernelGGL((inv_diag_kernel<T, 4>), dim3(N), dim3(64), 0, stream, S, P, (int)n, (int)N, err);
  else if (n <= 8)
    hipLaunchKernelGGL((inv_diag_kernel<T, 8>), dim3(N), dim3(64), 0, stream, S, P, (int)n, (int)N, err);
  else if (n <= 16)
    hipLaunchKernelGGL((inv_diag_kernel<T, 16>), dim3(N), dim3(64), 0, stream, S, P, (int)n, (int)N, err);
  else
    hipLaunchKernelGGL((inv_diag_kernel<T, 32>), dim3(N), dim3(64), 0, stream, S, P, (int)n, (int)N, err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t lds = 4 * (size_t)n * n * sizeof(T);
  hipLaunchKernelGGL(stair_kernel<T>, dim3(N), dim3(kAuxThreads), lds, stream, S, P, (int)n, (int)N, kind);
  return hipGetLastError();
}

hipError_t precond_launch(int dtype, uint32_t n, uint32_t N, const void *S, void *Pinv, int kind, int *err,
                          hipStream_t stream) {
  if (n <= 32) {
    if (dtype == GBDPCG_F64)
      return precond_small<double>(n, N, static_cast<const double *>(S), static_cast<double *>(Pinv), kind, err,
                                   stream);
    return precond_small<float>(n, N, static_cast<const float *>(S), static_cast<float *>(Pinv), kind, err, stream);
  }
  const size_t ts = dtype == GBDPCG_F64 ? 8 : 4;
  const size_t lds = (7 * (size_t)n * n + n) * ts;
  if (dtype == GBDPCG_F64) {
    hipFuncSetAttribute(reinterpret_cast<const void *>(&precond_kernel<double>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(precond_kernel<double>, dim3(N), dim3(kAuxThreads), lds, stream,
                       static_cast<const double *>(S), static_cast<double *>(Pinv), (int)n, (int)N, kind, err);
  } else {
    hipFuncSetAttribute(reinterpret_cast<const void *>(&precond_kernel<float>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(precond_kernel<float>, dim3(N), dim3(kAuxThreads), lds, stream,
                       static_cast<const float *>(S), static_cast<float *>(Pinv), (int)n, (int)N, kind, err);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// LQR Schur complement: block row k of S from (A, B, Q, R), one workgroup per knot
// ---------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(kAuxThreads) schur_kernel(const T *__restrict__ A, const T *__restrict__ B,
                                                           const T *__restrict__ Qd, const T *__restrict__ Rd,
                                                           T *__restrict__ S, int n, int m, int N) {
  const int k = blockIdx.x;
  const size_t bs = (size_t)3 * n * n;
  T *row = S + (size_t)k * bs;
  for (int e = threadIdx.x; e < 3 * n * n; e += blockDim.x) {
    const int col = e / n, r = e - col * n;
    const int seg = col / n, c = col - seg * n;
    T val = 0;
    if (seg == 0) {
      // phi_k = -A_{k-1} Q_{k-1}^-1
      if (k > 0) val = -A[(size_t)(k - 1) * n * n + r * n + c] / Qd[(size_t)(k - 1) * n + c];
    } else if (seg == 1) {
      if (k > 0) {
        const T *Ak = A + (size_t)(k - 1) * n * n;
        const T *Bk = B + (size_t)(k - 1) * n * m;
        const T *qk = Qd + (size_t)(k - 1) * n;
        const T *rk = Rd + (size_t)(k - 1) * m;
        for (int i = 0; i < n; ++i) val += Ak[r * n + i] * Ak[c * n + i] / qk[i];
        for (int j = 0; j < m; ++j) val += Bk[r * m + j] * Bk[c * m + j] / rk[j];
      }
      if (r == c) val += (T)1 / Qd[(size_t)k * n + r];
    } else {
      // (k,k+1) = phi_{k+1}^T: element (r,c) = -A_k[c][r] / q_k[r]
      if (k < N - 1) val = -A[(size_t)k * n * n + c * n + r] / Qd[(size_t)k * n + r];
    }
    row[e] = val;
  }
}

hipError_t schur_launch(int dtype, uint32_t n, uint32_t m, uint32_t N, const void *A, const void *B,
                        const void *Qd, const void *Rd, void *S, hipStream_t stream) {
  if (dtype == GBDPCG_F64) {
    hipLaunchKernelGGL(schur_kernel<double>, dim3(N), dim3(kAuxThreads), 0, stream, static_cast<const double *>(A),
                       static_cast<const double *>(B), static_cast<const double *>(Qd),
                       static_cast<const double *>(Rd), static_cast<double *>(S), (int)n, (int)m, (int)N);
  } else {
    hipLaunchKernelGGL(schur_kernel<float>, dim3(N), dim3(kAuxThreads), 0, stream, static_cast<const float *>(A),
                       static_cast<const float *>(B), static_cast<const float *>(Qd),
                       static_cast<const float *>(Rd), static_cast<float *>(S), (int)n, (int)m, (int)N);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// warm start: lambda[k] <- lambda[k+1] (k < N-1), gamma += noise
// ---------------------------------------------------------------------------

// One workgroup walks the vector in chunks: a chunk is read completely before
// any of it is written, and a later chunk only reads positions no earlier chunk
// has written (the shift moves values towards lower indices), so the in-place
// shift is race free.
template <typename T>
__device__ __forceinline__ void shift_knots(T *lambda, int n, int N) {
  const int limit = n * (N - 1);
  for (int base = 0; base < limit; base += kAuxThreads) {
    const int i = base + threadIdx.x;
    T v = 0;
    if (i < limit) v = lambda[i + n];
    __syncthreads();
    if (i < limit) lambda[i] = v;
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(kAuxThreads) warm_start_kernel(T *lambda, T *gamma, const T *noise, int n, int N) {
  shift_knots(lambda, n, N);
  if (noise != nullptr) {
    for (int i = threadIdx.x; i < n * N; i += kAuxThreads) gamma[i] += noise[i];
  }
}

// Batched form: workgroup b handles system b of a plan (forward systems only);
// the noise is generated in place from (seed, b, i), so the launch reads nothing
// but the plan's descriptors.
template <typename T>
__global__ void __launch_bounds__(kAuxThreads) batch_warm_start_kernel(const KSys *systems, uint64_t seed,
                                                                       double scale) {
  const KSys &d = systems[blockIdx.x];
  if (!(d.flags & KF_FWD)) return;
  const int n = (int)d.n, N = (int)d.N;
  shift_knots(static_cast<T *>(d.lambda), n, N);
  if (scale != 0.0) {
    T *gamma = static_cast<T *>(const_cast<void *>(d.gamma));
    const uint64_t key = counter_key(seed, blockIdx.x);
    for (int i = threadIdx.x; i < n * N; i += kAuxThreads) gamma[i] += (T)(scale * counter_normal(key, (uint64_t)i));
  }
}

hipError_t warm_start_launch(int dtype, uint32_t n, uint32_t N, void *lambda, void *gamma, const void *noise,
                             hipStream_t stream) {
  if (dtype == GBDPCG_F64) {
    hipLaunchKernelGGL(warm_start_kernel<double>, dim3(1), dim3(kAuxThreads), 0, stream,
                       static_cast<double *>(lambda), static_cast<double *>(gamma),
                       static_cast<const double *>(noise), (int)n, (int)N);
  } else {
    hipLaunchKernelGGL(warm_start_kernel<float>, dim3(1), dim3(kAuxThreads), 0, stream,
                       static_cast<float *>(lambda), static_cast<float *>(gamma), static_cast<const float *>(noise),
                       (int)n, (int)N);
  }
  return hipGetLastError();
}

hipError_t batch_warm_start_launch(int dtype, const KSys *systems, uint32_t count, uint64_t seed, double scale,
                                   hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (dtype == GBDPCG_F64)
    hipLaunchKernelGGL(batch_warm_start_kernel<double>, dim3(count), dim3(kAuxThreads), 0, stream, systems, seed,
                       scale);
  else
    hipLaunchKernelGGL(batch_warm_start_kernel<float>, dim3(count), dim3(kAuxThreads), 0, stream, systems, seed,
                       scale);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Exchange floor: rounds of the solvers' exchange alone (every rank publishes
// three words into one of two alternating records, wave 0 waits for every
// rank's words, then a workgroup barrier) -- the latency under one PCG
// iteration.  Round 0 checks the placement as the solvers do (all ranks on one
// XCD: the later rounds' stores stay in its L2).  Workgroup b is rank b/stride.
// ---------------------------------------------------------------------------
template <int NT>
__global__ void __launch_bounds__(NT) exchange_floor_kernel(uint64_t *slots, uint32_t G, uint32_t rounds,
                                                            uint32_t stride, uint32_t allow_xcd, int *out) {
  __shared__ int ctl[4];
  if (blockIdx.x % stride) return;
  const uint32_t rank = blockIdx.x / stride;
  if (rank >= G) return;
  const int t = (int)threadIdx.x;
  auto rec = [&](uint32_t type, uint32_t g) { return slots + ((size_t)type * G + g) * 4; };
  if (t == 0) {
    const uint32_t ea = (uint32_t)(xload(rec(0, rank)) >> 32);
    const uint32_t eb = (uint32_t)(xload(rec(1, rank)) >> 32);
    uint32_t base = (ea > eb ? ea : eb) + 1u;
    if (base > 0xFFF00000u) base = 1u;
    ctl[0] = 0;
    ctl[1] = (int)base;
  }
  __syncthreads();
  const uint32_t base = (uint32_t)ctl[1];
  bool xcd = false;
  for (uint32_t r = 0; r <= rounds; ++r) {
    const uint32_t epoch = base + r;
    const uint32_t type = r & 1u;
    if (t < 3) {
      const uint32_t x = 1u + (uint32_t)xcc_id();
      const uint32_t v = r == 0 ? (t == 0 ? 1u : t == 1 ? x : x * x) : rank + (uint32_t)t;
      xstore_sel(rec(type, rank) + t, v, epoch, xcd);
    }
    if (t < 64) {
      float s1 = 0.f, s2 = 0.f;  // sums of small integers: exact in float
      bool ok = true;
      for (uint32_t g = (uint32_t)t; g < G; g += 64) {
        uint32_t w[3];
        if (!xwait_words<3>(rec(type, g), epoch, w)) {
          ok = false;
          break;
        }
        s1 += (float)w[1];
        s2 += (float)w[2];
      }
      if (!ok) ctl[0] = 1;
      if (r == 0) {
        s1 = wave_sum(s1);
        s2 = wave_sum(s2);
        if (t == 0) ctl[2] = (allow_xcd && G > 1 && (double)s1 * s1 == (double)G * s2) ? 1 : 0;
      }
    }
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(ctl[0]) != 0) break;
    if (r == 0) xcd = __builtin_amdgcn_readfirstlane(ctl[2]) != 0;
  }
  if (t == 0) {
    if (ctl[0]) __hip_atomic_fetch_or(out, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (rank == 0 && xcd) __hip_atomic_fetch_or(out, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

namespace {
const void *floor_kernel(int nt) {
  switch (nt) {
    case 64: return reinterpret_cast<const void *>(&exchange_floor_kernel<64>);
    case 256: return reinterpret_cast<const void *>(&exchange_floor_kernel<256>);
    case 512: return reinterpret_cast<const void *>(&exchange_floor_kernel<512>);
    default: return nullptr;
  }
}
}  // namespace

hipError_t exchange_floor_occupancy(int nt, int *wgs_per_cu) {
  const void *k = floor_kernel(nt);
  if (!k) return hipErrorInvalidValue;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(wgs_per_cu, k, nt, 0);
}

hipError_t exchange_floor_launch(int nt, uint64_t *slots, uint32_t G, uint32_t rounds, uint32_t stride,
                                 uint32_t allow_xcd, int *out, hipStream_t stream) {
  const void *k = floor_kernel(nt);
  if (!k) return hipErrorInvalidValue;
  void *args[] = {&slots, &G, &rounds, &stride, &allow_xcd, &out};
  return hipLaunchKernel(k, dim3(G * stride), dim3(nt), args, 0, stream);
}

// ---------------------------------------------------------------------------
// Resource rates measured on streaming kernels, the peaks the solvers'
// rooflines are read against.  Every access is in bounds by construction and
// no kernel waits on another workgroup.
// ---------------------------------------------------------------------------
constexpr int kRateThreads = 256;

// FP32 FMA: eight independent chains per lane (the build emits no packed-FP32
// instructions, so each FMA is one lane-op).  The final store never happens for
// the arguments used (the chains converge to c / (1 - m)); it keeps them live.
__global__ void __launch_bounds__(kRateThreads) valu_rate_kernel(float m, float c, uint32_t reps, float *out) {
  float a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = (float)(threadIdx.x + i) * 1e-3f;
  for (uint32_t r = 0; r < reps; ++r) {
#pragma unroll
    for (int s = 0; s < 16; ++s)
#pragma unroll
      for (int i = 0; i < 8; ++i) a[i] = __builtin_fmaf(a[i], m, c);
  }
  const float t = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  if (t == -1.0f) out[blockIdx.x] = t;
}

// LDS reads: 16-byte loads of consecutive lanes (conflict-free), four per step
// from one of two alternating halves of a 32 KiB array; two three-input integer
// adds per load keep the VALU well below its own rate.
__global__ void __launch_bounds__(kRateThreads) lds_rate_kernel(uint32_t reps, uint32_t *out) {
  constexpr int kEntries = 8 * kRateThreads;  // uint4 entries: 32 KiB
  __shared__ uint4 buf[kEntries];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < (uint32_t)kEntries; i += kRateThreads) buf[i] = make_uint4(i, 3u * i, 5u * i, 7u * i);
  __syncthreads();
  uint32_t acc0 = 0, acc1 = 0, half = 0;
  for (uint32_t r = 0; r < reps; ++r) {
    // the empty asm makes the offset opaque, so the loads cannot be hoisted out
    // of the loop (the two halves would otherwise be loop-invariant)
    asm volatile("" : "+v"(half));
    const uint32_t base = t + half;  // < 8 * kRateThreads with s < 4 below
#pragma unroll
    for (uint32_t s = 0; s < 4; ++s) {
      const uint4 v = buf[base + s * kRateThreads];
      acc0 = acc0 + v.x + v.y;
      acc1 = acc1 + v.z + v.w;
    }
    half ^= 4u * kRateThreads;
  }
  if ((acc0 ^ acc1) == 0x9E3779B9u) out[blockIdx.x] = acc0;
}

// HBM reads: a grid-stride pass over `count` 16-byte words, four loads in flight
// per lane.
__global__ void __launch_bounds__(kRateThreads) hbm_rate_kernel(const uint4 *__restrict__ src, size_t count,
                                                                uint32_t *out) {
  const size_t step = (size_t)gridDim.x * kRateThreads;
  uint32_t acc0 = 0, acc1 = 0;
  size_t i = (size_t)blockIdx.x * kRateThreads + threadIdx.x;
  for (; i + 3 * step < count; i += 4 * step) {
    const uint4 v0 = src[i], v1 = src[i + step], v2 = src[i + 2 * step], v3 = src[i + 3 * step];
    acc0 = acc0 + v0.x + v0.y + v1.x + v1.y + v2.x + v2.y + v3.x + v3.y;
    acc1 = acc1 + v0.z + v0.w + v1.z + v1.w + v2.z + v2.w + v3.z + v3.w;
  }
  for (; i < count; i += step) {
    const uint4 v = src[i];
    acc0 = acc0 + v.x + v.y;
    acc1 = acc1 + v.z + v.w;
  }
  if ((acc0 ^ acc1) == 0x9E3779B9u) out[blockIdx.x] = acc0;
}

hipError_t rate_launch(int which, uint32_t grid, uint32_t reps, const void *src, size_t count, void *out,
                       hipStream_t stream) {
  switch (which) {
    case 0:
      hipLaunchKernelGGL(valu_rate_kernel, dim3(grid), dim3(kRateThreads), 0, stream, 0.5f, 0.25f, reps,
                         static_cast<float *>(out));
      break;
    case 1:
      hipLaunchKernelGGL(lds_rate_kernel, dim3(grid), dim3(kRateThreads), 0, stream, reps,
                         static_cast<uint32_t *>(out));
      break;
    case 2:
      hipLaunchKernelGGL(hbm_rate_kernel, dim3(grid), dim3(kRateThreads), 0, stream,
                         static_cast<const uint4 *>(src), count, static_cast<uint32_t *>(out));
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gbdpcg
