"""CPU oracle for the block-tridiagonal PCG path — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker (or the timed CPU baseline).  The product path
(gbd-pcg_amd/) never imports it.

Contents
  * ctypes bindings of liboracle.so (pcg_oracle.c): the reference PCG iteration
    (include/pcg.cuh:118-208) in FP64 and FP32, its Chronopoulos-Gear form, the
    block-row matvec (include/utils.cuh:46-85) and a block-tridiagonal solve;
  * numpy restatements of the cbtd layout (include/pcg.cuh:103-110,
    include/utils.cuh:58-84): cbtd <-> dense;
  * the block-Jacobi and symmetric-stair preconditioners (definition only);
  * the dense FP64 adjoint: g_gamma = S^-1 g_lambda,
    g_S = -1/2 (g_gamma lambda^T + lambda g_gamma^T) on the block-tridiagonal pattern.

Parity pins of every function are in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")

EXIT_ABS = 0
EXIT_REL = 1
CONVERGED, MAX_ITER, BREAKDOWN, NONFINITE = 0, 1, 2, 3

_lib = None


def lib():
    global _lib
    if _lib is None:
        _ensure_built()
        _lib = ctypes.CDLL(LIBPATH)
        for sfx, ct in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            P = ctypes.POINTER(ct)
            for name in ("oracle_pcg", "oracle_pcg_fused"):
                f = getattr(_lib, f"{name}_{sfx}")
                f.restype = ctypes.c_int
                f.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
            f = getattr(_lib, f"oracle_bdmv_{sfx}")
            f.restype = None
            f.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P]
        f = _lib.oracle_block_solve_f64
        f.restype = ctypes.c_int
        PD = ctypes.POINTER(ctypes.c_double)
        f.argtypes = [ctypes.c_int, ctypes.c_int, PD, PD, PD]
    return _lib


def _ensure_built():
    if not os.path.exists(LIBPATH):
        import importlib.util
        spec = importlib.util.spec_from_file_location("oracle_build", os.path.join(HERE, "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()


def _ptr(a: np.ndarray):
    ct = ctypes.c_double if a.dtype == np.float64 else ctypes.c_float
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _sfx(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def pcg(S, Pinv, gamma, lam0, n, N, tol, max_iter, mode=EXIT_ABS, dtype=np.float64, fused=False):
    """Runs the reference iteration (or its fused form) on the CPU.

    Returns (lambda, iters, status, eta_final, rr_final, eta_history)."""
    _ensure_built()
    dt = np.dtype(dtype)
    S = np.ascontiguousarray(S, dtype=dt)
    P = np.ascontiguousarray(Pinv, dtype=dt)
    g = np.ascontiguousarray(gamma, dtype=dt)
    lam = np.array(lam0, dtype=dt, copy=True)
    hist = np.zeros(max_iter + 1, dtype=np.float64)
    st = ctypes.c_int(0)
    eta = ctypes.c_double(0)
    rr = ctypes.c_double(0)
    f = getattr(lib(), ("oracle_pcg_fused_" if fused else "oracle_pcg_") + _sfx(dt))
    it = f(n, N, _ptr(S), _ptr(P), _ptr(g), _ptr(lam), float(tol), int(max_iter), int(mode),
           hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(st), ctypes.byref(eta), ctypes.byref(rr))
    return lam, it, st.value, eta.value, rr.value, hist[: it + 1]


def bdmv(A, x, n, N, dtype=np.float64):
    _ensure_built()
    dt = np.dtype(dtype)
    A = np.ascontiguousarray(A, dtype=dt)
    x = np.ascontiguousarray(x, dtype=dt)
    y = np.zeros(n * N, dtype=dt)
    getattr(lib(), "oracle_bdmv_" + _sfx(dt))(n, N, _ptr(A), _ptr(x), _ptr(y))
    return y


def block_solve(S, b, n, N):
    """FP64 block-tridiagonal direct solve S x = b (block LU)."""
    _ensure_built()
    S = np.ascontiguousarray(S, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n * N)
    rc = lib().oracle_block_solve_f64(n, N, _ptr(S), _ptr(b), _ptr(x))
    if rc != 0:
        raise np.linalg.LinAlgError("singular diagonal block in block LU")
    return x


# ---------------------------------------------------------------------------
# cbtd layout (numpy)
# ---------------------------------------------------------------------------

def cbtd_to_dense(A, n, N):
    """Dense (nN x nN) matrix of a cbtd array.  Block row k is n x 3n column-major,
    holding blocks (k,k-1), (k,k), (k,k+1); the two padding blocks are dropped."""
    A = np.asarray(A, dtype=np.float64).reshape(N, 3, n, n)  # [k, seg, col, row]
    M = np.zeros((n * N, n * N))
    for k in range(N):
        for seg in range(3):
            j = k - 1 + seg
            if 0 <= j < N:
                M[k * n:(k + 1) * n, j * n:(j + 1) * n] = A[k, seg].T
    return M


def dense_to_cbtd(M, n, N):
    """cbtd array of the block-tridiagonal part of a dense matrix (pads zero)."""
    M = np.asarray(M, dtype=np.float64)
    A = np.zeros((N, 3, n, n))
    for k in range(N):
        for seg in range(3):
            j = k - 1 + seg
            if 0 <= j < N:
                A[k, seg] = M[k * n:(k + 1) * n, j * n:(j + 1) * n].T
    return A.reshape(-1)


def precond(S, n, N, kind):
    """Block-Jacobi (kind 1) or symmetric-stair (kind 2) preconditioner of a cbtd S,
    from its definition: theta_k = S_kk, phi_k = S_{k,k-1};
    Jacobi: blockdiag(theta_k^-1);
    SS: diag theta_k^-1, (k,k-1) block -theta_k^-1 phi_k theta_{k-1}^-1,
        (k-1,k) its transpose."""
    D = cbtd_to_dense(S, n, N)
    P = np.zeros_like(D)
    inv = [np.linalg.inv(D[k * n:(k + 1) * n, k * n:(k + 1) * n]) for k in range(N)]
    for k in range(N):
        P[k * n:(k + 1) * n, k * n:(k + 1) * n] = inv[k]
        if kind == 2 and k > 0:
            phi = D[k * n:(k + 1) * n, (k - 1) * n:k * n]
            B = -inv[k] @ phi @ inv[k - 1]
            P[k * n:(k + 1) * n, (k - 1) * n:k * n] = B
            P[(k - 1) * n:k * n, k * n:(k + 1) * n] = B.T
    return dense_to_cbtd(P, n, N)


def pattern_mask(n, N):
    """cbtd array with ones on stored (non-padding) entries."""
    return dense_to_cbtd(cbtd_to_dense(np.ones(3 * n * n * N), n, N) != 0, n, N)


def adjoint_dense(S, lam, glam, n, N):
    """FP64 dense adjoint of lambda = S^-1 gamma for symmetric S:
    g_gamma = S^-1 g_lambda; g_S = -1/2 (g_gamma lambda^T + lambda g_gamma^T)
    restricted to the block-tridiagonal pattern (cbtd)."""
    D = cbtd_to_dense(S, n, N)
    gg = np.linalg.solve(D, np.asarray(glam, dtype=np.float64))
    lam = np.asarray(lam, dtype=np.float64)
    G = -0.5 * (np.outer(gg, lam) + np.outer(lam, gg))
    return gg, dense_to_cbtd(G, n, N)


def dense_solve(S, gamma, n, N):
    return np.linalg.solve(cbtd_to_dense(S, n, N), np.asarray(gamma, dtype=np.float64))
