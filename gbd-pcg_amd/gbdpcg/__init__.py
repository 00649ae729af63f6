"""Python binding of libgbdpcg.so — the MI355X block-tridiagonal PCG solver.

The compute path is the HIP library; this module only marshals arguments through
the C ABI declared in include/gbdpcg.h (ctypes).  PyTorch is used for device
memory and streams (tensor.data_ptr(), torch.cuda.current_stream()).  There is
no CPU fallback: if the library is missing, importing the binding raises.

Reference interface mirrored (file:line in the reference tree):
  solve_host  <- solvePCG(h_S, h_Pinv, h_gamma, h_lambda, n, N, cfg)  include/interface.cuh:24-91
  solve       <- solvePCG(n, N, d_S, d_Pinv, d_gamma, d_lambda, ...)  include/interface.cuh:94-146
  Config      <- pcg_config<T>                                       include/types.cuh:18-35
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libgbdpcg.so")

# ---- constants (include/gbdpcg.h) ----
OK = 0
ERR_INVALID, ERR_TOO_LARGE, ERR_DEVICE, ERR_TIMEOUT, ERR_NOMEM, ERR_UNSUPPORTED = -1, -2, -3, -4, -5, -6
CONVERGED, MAX_ITER, BREAKDOWN, NONFINITE, TIMED_OUT = 0, 1, 2, 3, 4
F32, F64 = 0, 1
EXIT_ABS_ETA, EXIT_REL_RESID = 0, 1
PRECOND_JACOBI, PRECOND_SS = 1, 2
RECUR_REFERENCE, RECUR_FUSED = 0, 1
KERNEL_AUTO, KERNEL_LDS, KERNEL_REGISTER, KERNEL_ROWCAST = 0, 1, 2, 3
EXCHANGE_AUTO, EXCHANGE_DEVICE = 0, 1
SYS_FORWARD, SYS_ADJOINT = 1, 2

_ERRNAMES = {ERR_INVALID: "INVALID", ERR_TOO_LARGE: "TOO_LARGE", ERR_DEVICE: "DEVICE", ERR_TIMEOUT: "TIMEOUT",
             ERR_NOMEM: "NOMEM", ERR_UNSUPPORTED: "UNSUPPORTED"}


class GbdpcgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gbdpcg error {code} ({_ERRNAMES.get(code, '?')}): {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("exit_tol", ctypes.c_double), ("max_iter", ctypes.c_uint32), ("grid", ctypes.c_uint32 * 3),
                ("block", ctypes.c_uint32 * 3), ("empty_pinv", ctypes.c_int32), ("exit_mode", ctypes.c_int32),
                ("precond", ctypes.c_int32), ("recurrence", ctypes.c_int32), ("stream", ctypes.c_void_p),
                ("kernel", ctypes.c_uint32), ("knots_per_wg", ctypes.c_uint32), ("threads", ctypes.c_uint32),
                ("batch_split", ctypes.c_uint32), ("exchange_scope", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 3)]


class Result(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int32), ("status", ctypes.c_int32), ("adj_iters", ctypes.c_int32),
                ("adj_status", ctypes.c_int32), ("eta", ctypes.c_double), ("rr", ctypes.c_double),
                ("adj_eta", ctypes.c_double), ("adj_rr", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SystemDesc(ctypes.Structure):
    _fields_ = [("S", ctypes.c_void_p), ("Pinv", ctypes.c_void_p), ("gamma", ctypes.c_void_p),
                ("lam", ctypes.c_void_p), ("glambda", ctypes.c_void_p), ("ggamma", ctypes.c_void_p),
                ("gS", ctypes.c_void_p), ("n", ctypes.c_uint32), ("N", ctypes.c_uint32),
                ("exit_tol", ctypes.c_double), ("max_iter", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("wgs", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("grid", ctypes.c_uint32), ("rounds", ctypes.c_uint32), ("wgs_per_cu", ctypes.c_uint32),
                ("lds_bytes", ctypes.c_uint32), ("systems", ctypes.c_uint32), ("total_items", ctypes.c_uint32),
                ("est_cost", ctypes.c_double), ("launches", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class TuneReport(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_uint32), ("from_cache", ctypes.c_uint32), ("knots_per_wg", ctypes.c_uint32),
                ("threads", ctypes.c_uint32), ("wgs", ctypes.c_uint32), ("kernel", ctypes.c_uint32),
                ("best_us", ctypes.c_double), ("default_us", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


RESULT_BYTES = ctypes.sizeof(Result)
NOALLOC_SCRATCH_BYTES = 1 << 20  # GBDPCG_NOALLOC_SCRATCH_BYTES

_lib = None

# (name, restype, argtypes) of every exported symbol of include/gbdpcg.h
_vp, _u32, _i32, _u64, _dbl, _sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64, \
    ctypes.c_double, ctypes.c_size_t
_PC, _PR, _PU = ctypes.POINTER(Config), ctypes.POINTER(Result), ctypes.POINTER(ctypes.c_uint32)
_PD = ctypes.POINTER(ctypes.c_double)
SIGNATURES = [
    ("gbdpcg_config_init", None, [_PC]),
    ("gbdpcg_version", ctypes.c_char_p, []),
    ("gbdpcg_last_error", ctypes.c_char_p, []),
    ("gbdpcg_lds_bytes", _sz, [_i32, _u32, _u32, _u32]),
    ("gbdpcg_check_occupancy", _i32, [_i32, _u32, _u32, _PU]),
    ("gbdpcg_solve_f32", _i32, [_vp, _vp, _vp, _vp, _u32, _u32, _PC]),
    ("gbdpcg_solve_f64", _i32, [_vp, _vp, _vp, _vp, _u32, _u32, _PC]),
    ("gbdpcg_solve_device_f32", _i32, [_u32, _u32, _vp, _vp, _vp, _vp, _PC, _PR]),
    ("gbdpcg_solve_device_f64", _i32, [_u32, _u32, _vp, _vp, _vp, _vp, _PC, _PR]),
    ("gbdpcg_solve_device_async_f32", _i32, [_u32, _u32, _vp, _vp, _vp, _vp, _PC, _vp]),
    ("gbdpcg_solve_device_async_f64", _i32, [_u32, _u32, _vp, _vp, _vp, _vp, _PC, _vp]),
    ("gbdpcg_solve_device_noalloc_f32", _i32, [_u32, _u32, _vp, _vp, _vp, _vp, _PC, _vp]),
    ("gbdpcg_solve_device_noalloc_f64", _i32, [_u32, _u32, _vp, _vp, _vp, _vp, _PC, _vp]),
    ("gbdpcg_solve_csr_f32", _i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _PC]),
    ("gbdpcg_solve_csr_f64", _i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _PC]),
    ("gbdpcg_csr_to_cbtd", _i32, [_i32, _vp, _vp, _vp, _u32, _u32, _vp]),
    ("gbdpcg_adjoint_device_f32", _i32, [_u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _PC, _PR]),
    ("gbdpcg_adjoint_device_f64", _i32, [_u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _PC, _PR]),
    ("gbdpcg_batch_create", _i32, [ctypes.POINTER(_vp), _i32, ctypes.POINTER(SystemDesc), _u32, _PC, _u32]),
    ("gbdpcg_batch_run", _i32, [_vp, _vp]),
    ("gbdpcg_batch_results_device", _vp, [_vp]),
    ("gbdpcg_batch_results", _i32, [_vp, _PR, _vp]),
    ("gbdpcg_batch_info_get", _i32, [_vp, ctypes.POINTER(BatchInfo)]),
    ("gbdpcg_batch_wgs", _i32, [_vp, _PU]),
    ("gbdpcg_batch_destroy", _i32, [_vp]),
    ("gbdpcg_batch_warm_start", _i32, [_vp, _u64, _dbl, _vp]),
    ("gbdpcg_tune_single", _i32, [_i32, _u32, _u32, _vp, _vp, _vp, _PC, _u32, ctypes.POINTER(TuneReport)]),
    ("gbdpcg_batch_autotune", _i32, [ctypes.POINTER(_vp), _i32, ctypes.POINTER(SystemDesc), _u32, _PC, _u32,
                                     ctypes.POINTER(TuneReport)]),
    ("gbdpcg_precond_device", _i32, [_i32, _u32, _u32, _vp, _vp, _i32, _vp]),
    ("gbdpcg_precond_host", _i32, [_i32, _u32, _u32, _vp, _vp, _i32]),
    ("gbdpcg_lqr_generate", _i32, [_i32, _u32, _u32, _u32, _u64, _dbl, _vp, _vp, _PD, _PD, _PD, _PD]),
    ("gbdpcg_lqr_schur_host", _i32, [_i32, _u32, _u32, _u32, _PD, _PD, _PD, _PD, _vp]),
    ("gbdpcg_lqr_schur_device", _i32, [_i32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("gbdpcg_randn", _i32, [_i32, _u64, _u64, _dbl, _vp]),
    ("gbdpcg_counter_randn", _i32, [_i32, _u64, _u64, _u64, _dbl, _vp]),
    ("gbdpcg_exchange_floor", _i32, [_u32, _u32, _u32, ctypes.POINTER(_dbl), ctypes.POINTER(_u32)]),
    ("gbdpcg_measure_rates", _i32, [ctypes.POINTER(_dbl)] * 3),
    ("gbdpcg_warm_start_device", _i32, [_i32, _u32, _u32, _vp, _vp, _vp, _vp]),
]


def lib():
    """Loads libgbdpcg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is missing: build it with `python gbd-pcg_amd/build.py`")
        L = ctypes.CDLL(LIBPATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().gbdpcg_last_error().decode()


def check(rc: int) -> int:
    if rc < 0:
        raise GbdpcgError(rc, last_error())
    return rc


def version() -> str:
    return lib().gbdpcg_version().decode()


def _dtype_code(dt) -> int:
    s = str(dt)
    if not s.startswith("torch."):
        s = str(np.dtype(dt))
    if s in ("torch.float64", "float64"):
        return F64
    if s in ("torch.float32", "float32"):
        return F32
    raise TypeError(f"unsupported dtype {dt}")


def make_config(tol: float = 1e-6, max_iter: int = 25, exit_mode: int = EXIT_ABS_ETA,
                recurrence: int = RECUR_REFERENCE, precond: int = PRECOND_SS, wgs: int = 0,
                stream: Optional[int] = None, kernel: int = KERNEL_AUTO, knots_per_wg: int = 0,
                threads: int = 0, batch_split: bool = False, exchange_scope: int = EXCHANGE_AUTO) -> Config:
    cfg = Config()
    lib().gbdpcg_config_init(ctypes.byref(cfg))
    cfg.exit_tol = float(tol)
    cfg.max_iter = int(max_iter)
    cfg.exit_mode = int(exit_mode)
    cfg.recurrence = int(recurrence)
    cfg.precond = int(precond)
    cfg.grid[0] = int(wgs)
    cfg.stream = stream
    cfg.kernel = int(kernel)
    cfg.knots_per_wg = int(knots_per_wg)
    cfg.threads = int(threads)
    cfg.batch_split = 1 if batch_split else 0
    cfg.exchange_scope = int(exchange_scope)
    return cfg


def _stream_handle(stream) -> Optional[int]:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream or None
    if isinstance(stream, int):
        return stream or None
    return stream.cuda_stream or None


def _dev(t, name):
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous device tensor")
    return t.data_ptr()


def _check_system(S, n: int, N: int, mats=(), vecs=(), where: str = "") -> int:
    """Checks one system's device tensors before anything is launched: every
    tensor (given as (name, tensor) pairs; None entries are skipped) is a
    contiguous device tensor on S's device with S's dtype, S and the other cbtd
    arrays hold 3 n^2 N elements and the vectors n N.  Returns the dtype code.
    A short or mistyped buffer would otherwise be read or written out of bounds
    by the kernels."""
    if int(n) <= 0 or int(N) <= 0:
        raise ValueError(f"{where}n and N must be positive (got n={n}, N={N})")
    dt = _dtype_code(S.dtype)
    want = {"mat": 3 * int(n) * int(n) * int(N), "vec": int(n) * int(N)}
    for kind, group in (("mat", (("S", S),) + tuple(mats)), ("vec", tuple(vecs))):
        for name, t in group:
            if t is None:
                continue
            if t.dtype != S.dtype:
                raise ValueError(f"{where}{name} has dtype {t.dtype}, S has {S.dtype}")
            if t.numel() != want[kind]:
                raise ValueError(f"{where}{name} has {t.numel()} elements, expected {want[kind]} "
                                 f"({'3 n^2 N' if kind == 'mat' else 'n N'} with n={n}, N={N})")
            if t.device != S.device:
                raise ValueError(f"{where}{name} is on {t.device}, S is on {S.device}")
            _dev(t, where + name)
    return dt


def _check_result_buffer(d_result):
    if d_result is None:
        return
    _dev(d_result, "d_result")
    if d_result.numel() * d_result.element_size() < RESULT_BYTES:
        raise ValueError(f"d_result must hold at least {RESULT_BYTES} bytes")


# ---------------------------------------------------------------------------
# single-system solves (torch device tensors)
# ---------------------------------------------------------------------------

def solve(S, gamma, lam, n: int, N: int, Pinv=None, *, tol=1e-6, max_iter=25, exit_mode=EXIT_ABS_ETA,
          recurrence=RECUR_REFERENCE, precond=PRECOND_SS, wgs=0, stream=None, kernel=KERNEL_AUTO,
          threads=0, exchange_scope=EXCHANGE_AUTO) -> Result:
    """Solves S lam = gamma in place (lam holds the initial guess).  Waits for the
    stream; returns the result record."""
    dt = _check_system(S, n, N, mats=(("Pinv", Pinv),), vecs=(("gamma", gamma), ("lam", lam)))
    cfg = make_config(tol, max_iter, exit_mode, recurrence, precond, wgs, _stream_handle(stream), kernel,
                      threads=threads, exchange_scope=exchange_scope)
    res = Result()
    f = lib().gbdpcg_solve_device_f64 if dt == F64 else lib().gbdpcg_solve_device_f32
    check(f(n, N, _dev(S, "S"), _dev(Pinv, "Pinv"), _dev(gamma, "gamma"), _dev(lam, "lam"), ctypes.byref(cfg),
            ctypes.byref(res)))
    return res


def solve_async(S, gamma, lam, n: int, N: int, Pinv=None, d_result=None, *, tol=1e-6, max_iter=25,
                exit_mode=EXIT_ABS_ETA, recurrence=RECUR_REFERENCE, precond=PRECOND_SS, wgs=0, stream=None,
                kernel=KERNEL_AUTO, threads=0, exchange_scope=EXCHANGE_AUTO) -> None:
    """Launches a solve without waiting; the result record goes to d_result (a
    device uint8 tensor of RESULT_BYTES bytes) when given."""
    dt = _check_system(S, n, N, mats=(("Pinv", Pinv),), vecs=(("gamma", gamma), ("lam", lam)))
    _check_result_buffer(d_result)
    cfg = make_config(tol, max_iter, exit_mode, recurrence, precond, wgs, _stream_handle(stream), kernel,
                      threads=threads, exchange_scope=exchange_scope)
    f = lib().gbdpcg_solve_device_async_f64 if dt == F64 else lib().gbdpcg_solve_device_async_f32
    check(f(n, N, _dev(S, "S"), _dev(Pinv, "Pinv"), _dev(gamma, "gamma"), _dev(lam, "lam"), ctypes.byref(cfg),
            _dev(d_result, "d_result")))


def solve_noalloc(S, gamma, lam, n: int, N: int, Pinv, d_result=None, *, tol=1e-6, max_iter=25,
                  exit_mode=EXIT_ABS_ETA, recurrence=RECUR_REFERENCE, wgs=0, stream=None, kernel=KERNEL_AUTO,
                  threads=0, exchange_scope=EXCHANGE_AUTO) -> None:
    """solve_async without any device allocation: the exchange records live in the
    library's fixed per-device scratch (NOALLOC_SCRATCH_BYTES).  Pinv is
    required.  Calls that overlap on one device race on that scratch."""
    if Pinv is None:
        raise ValueError("solve_noalloc needs Pinv")
    dt = _check_system(S, n, N, mats=(("Pinv", Pinv),), vecs=(("gamma", gamma), ("lam", lam)))
    _check_result_buffer(d_result)
    cfg = make_config(tol, max_iter, exit_mode, recurrence, PRECOND_SS, wgs, _stream_handle(stream), kernel,
                      threads=threads, exchange_scope=exchange_scope)
    f = lib().gbdpcg_solve_device_noalloc_f64 if dt == F64 else lib().gbdpcg_solve_device_noalloc_f32
    check(f(n, N, _dev(S, "S"), _dev(Pinv, "Pinv"), _dev(gamma, "gamma"), _dev(lam, "lam"), ctypes.byref(cfg),
            _dev(d_result, "d_result")))


def adjoint(S, gamma, lam, glam, ggamma, gS, n: int, N: int, Pinv, *, forward=True, tol=1e-6, max_iter=25,
            exit_mode=EXIT_ABS_ETA, recurrence=RECUR_REFERENCE, wgs=0, stream=None, kernel=KERNEL_AUTO,
            threads=0, exchange_scope=EXCHANGE_AUTO) -> Result:
    """(Forward +) adjoint solve; writes ggamma and gS (cbtd).  Waits for the stream."""
    if Pinv is None or glam is None or ggamma is None or gS is None or (forward and gamma is None):
        raise ValueError("adjoint needs Pinv, glam, ggamma, gS (and gamma when forward)")
    dt = _check_system(S, n, N, mats=(("Pinv", Pinv), ("gS", gS)),
                       vecs=(("gamma", gamma if forward else None), ("lam", lam), ("glam", glam),
                             ("ggamma", ggamma)))
    cfg = make_config(tol, max_iter, exit_mode, recurrence, PRECOND_SS, wgs, _stream_handle(stream), kernel,
                      threads=threads, exchange_scope=exchange_scope)
    res = Result()
    f = lib().gbdpcg_adjoint_device_f64 if dt == F64 else lib().gbdpcg_adjoint_device_f32
    check(f(n, N, _dev(S, "S"), _dev(Pinv, "Pinv"), _dev(gamma, "gamma") if forward else None, _dev(lam, "lam"),
            _dev(glam, "glam"), _dev(ggamma, "ggamma"), _dev(gS, "gS"), 1 if forward else 0, ctypes.byref(cfg),
            ctypes.byref(res)))
    return res


def precond_device(S, Pinv, n: int, N: int, kind: int = PRECOND_SS, stream=None) -> None:
    _check_system(S, n, N, mats=(("Pinv", Pinv),))
    check(lib().gbdpcg_precond_device(_dtype_code(S.dtype), n, N, _dev(S, "S"), _dev(Pinv, "Pinv"), kind,
                                      _stream_handle(stream)))


def lqr_schur_device(A, B, Qd, Rd, S, n: int, m: int, N: int, stream=None) -> None:
    check(lib().gbdpcg_lqr_schur_device(_dtype_code(S.dtype), n, m, N, _dev(A, "A"), _dev(B, "B") if m else None,
                                        _dev(Qd, "Qd"), _dev(Rd, "Rd") if m else None, _dev(S, "S"),
                                        _stream_handle(stream)))


def warm_start_device(lam, gamma, n: int, N: int, noise=None, stream=None) -> None:
    for name, t in (("lam", lam), ("gamma", gamma), ("noise", noise)):
        if t is not None and (t.dtype != lam.dtype or t.numel() != n * N):
            raise ValueError(f"{name} must have lam's dtype and n N = {n * N} elements")
        _dev(t, name)
    check(lib().gbdpcg_warm_start_device(_dtype_code(lam.dtype), n, N, _dev(lam, "lam"), _dev(gamma, "gamma"),
                                         _dev(noise, "noise"), _stream_handle(stream)))


# ---------------------------------------------------------------------------
# host-pointer solves (numpy)
# ---------------------------------------------------------------------------

def _np_ptr(a):
    return None if a is None else a.ctypes.data


def solve_host(S: np.ndarray, gamma: np.ndarray, lam: np.ndarray, n: int, N: int, Pinv: Optional[np.ndarray] = None,
               *, tol=1e-6, max_iter=25, exit_mode=EXIT_ABS_ETA, recurrence=RECUR_REFERENCE,
               precond=PRECOND_SS, wgs=0) -> int:
    """The reference's host-pointer call: lam (numpy, in/out).  Returns iterations."""
    dt = _dtype_code(S.dtype)
    if int(n) <= 0 or int(N) <= 0:
        raise ValueError(f"n and N must be positive (got n={n}, N={N})")
    for name, a, size in (("S", S, 3 * n * n * N), ("gamma", gamma, n * N), ("lam", lam, n * N),
                          ("Pinv", Pinv, 3 * n * n * N)):
        if a is None:
            continue
        if not a.flags["C_CONTIGUOUS"] or a.dtype != S.dtype:
            raise ValueError("arrays must be C-contiguous and share S's dtype")
        if a.size != size:
            raise ValueError(f"{name} has {a.size} elements, expected {size} (n={n}, N={N})")
    if not lam.flags["WRITEABLE"]:
        raise ValueError("lam is written in place: it must be writeable")
    cfg = make_config(tol, max_iter, exit_mode, recurrence, precond, wgs, None)
    f = lib().gbdpcg_solve_f64 if dt == F64 else lib().gbdpcg_solve_f32
    return check(f(_np_ptr(S), _np_ptr(Pinv), _np_ptr(gamma), _np_ptr(lam), n, N, ctypes.byref(cfg)))


def solve_csr(S_csr, gamma: np.ndarray, lam: np.ndarray, n: int, N: int, P_csr=None, *, tol=1e-6, max_iter=25,
              exit_mode=EXIT_ABS_ETA, recurrence=RECUR_REFERENCE, precond=PRECOND_SS, wgs=0) -> int:
    """CSR overload (reference include/interface.cuh:8-20): S_csr / P_csr are
    (row_ptr, col_ind, val) triples of an nN x nN matrix; lam (numpy) in/out."""
    def parts(m):
        if m is None:
            return None, None, None
        rp = np.ascontiguousarray(m[0], dtype=np.uint32)
        ci = np.ascontiguousarray(m[1], dtype=np.uint32)
        v = np.ascontiguousarray(m[2], dtype=lam.dtype)
        return rp, ci, v
    sr, sc, sv = parts(S_csr)
    pr, pc, pv = parts(P_csr)
    gamma = np.ascontiguousarray(gamma, dtype=lam.dtype)
    if not lam.flags["C_CONTIGUOUS"]:
        raise ValueError("lam must be C-contiguous")
    dt = _dtype_code(lam.dtype)
    cfg = make_config(tol, max_iter, exit_mode, recurrence, precond, wgs, None)
    f = lib().gbdpcg_solve_csr_f64 if dt == F64 else lib().gbdpcg_solve_csr_f32
    return check(f(_np_ptr(sr), _np_ptr(sc), _np_ptr(sv), _np_ptr(pr), _np_ptr(pc), _np_ptr(pv), _np_ptr(gamma),
                   _np_ptr(lam), n, N, ctypes.byref(cfg)))


def precond_host(S: np.ndarray, n: int, N: int, kind: int = PRECOND_SS) -> np.ndarray:
    P = np.zeros_like(S)
    check(lib().gbdpcg_precond_host(_dtype_code(S.dtype), n, N, _np_ptr(S), _np_ptr(P), kind))
    return P


def csr_to_cbtd(row_ptr, col_ind, val, n: int, N: int) -> np.ndarray:
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col_ind = np.ascontiguousarray(col_ind, dtype=np.uint32)
    val = np.ascontiguousarray(val)
    out = np.zeros(3 * n * n * N, dtype=val.dtype)
    check(lib().gbdpcg_csr_to_cbtd(_dtype_code(val.dtype), _np_ptr(row_ptr), _np_ptr(col_ind), _np_ptr(val), n, N,
                                   _np_ptr(out)))
    return out


def lqr_generate(n: int, m: int, N: int, seed: int, dtype=np.float32, dt: float = 0.1, with_lqr: bool = False):
    """Seeded LQR Schur-complement problem: returns (S cbtd, gamma) [+ (A, B, Qd, Rd)]."""
    dtp = np.dtype(dtype)
    S = np.zeros(3 * n * n * N, dtype=dtp)
    g = np.zeros(n * N, dtype=dtp)
    A = np.zeros(N * n * n)
    B = np.zeros(max(N * n * m, 1))
    Q = np.zeros(N * n)
    R = np.zeros(max(N * m, 1))
    P = ctypes.POINTER(ctypes.c_double)
    check(lib().gbdpcg_lqr_generate(_dtype_code(dtp), n, m, N, seed, dt, _np_ptr(S), _np_ptr(g),
                                    A.ctypes.data_as(P), B.ctypes.data_as(P), Q.ctypes.data_as(P),
                                    R.ctypes.data_as(P)))
    if with_lqr:
        return S, g, (A.reshape(N, n, n), B[: N * n * m].reshape(N, n, m), Q.reshape(N, n), R[: N * m].reshape(N, m))
    return S, g


def randn(count: int, seed: int, scale: float = 1.0, dtype=np.float32) -> np.ndarray:
    out = np.zeros(count, dtype=dtype)
    check(lib().gbdpcg_randn(_dtype_code(np.dtype(dtype)), seed, count, scale, _np_ptr(out)))
    return out


def counter_randn(count: int, seed: int, stream: int = 0, scale: float = 1.0, dtype=np.float32) -> np.ndarray:
    """The counter-based normals the batched warm start adds (same values on the host)."""
    out = np.zeros(count, dtype=dtype)
    check(lib().gbdpcg_counter_randn(_dtype_code(np.dtype(dtype)), seed, stream, count, scale, _np_ptr(out)))
    return out


def tune_single(S, gamma, n: int, N: int, Pinv=None, *, tol=1e-6, max_iter=25, exit_mode=EXIT_ABS_ETA,
                recurrence=RECUR_REFERENCE, precond=PRECOND_SS, kernel=KERNEL_AUTO, trials=5, stream=None) -> dict:
    """Times the candidate layouts of (n, N) on these inputs and makes the fastest
    the default for later automatic solves of this shape.  Returns the report."""
    _check_system(S, n, N, mats=(("Pinv", Pinv),), vecs=(("gamma", gamma),))
    cfg = make_config(tol, max_iter, exit_mode, recurrence, precond, 0, _stream_handle(stream), kernel)
    rep = TuneReport()
    check(lib().gbdpcg_tune_single(_dtype_code(S.dtype), n, N, _dev(S, "S"), _dev(Pinv, "Pinv"),
                                   _dev(gamma, "gamma"), ctypes.byref(cfg), int(trials), ctypes.byref(rep)))
    return rep.as_dict()


def exchange_floor(wgs: int, threads: int = 256, rounds: int = 2000) -> dict:
    """Time per round of the solvers' exchange alone (wgs workgroups placed as a
    system's ranks): the latency under one PCG iteration."""
    us, xcd = ctypes.c_double(), ctypes.c_uint32()
    check(lib().gbdpcg_exchange_floor(int(wgs), int(threads), int(rounds), ctypes.byref(us), ctypes.byref(xcd)))
    return {"us_per_round": us.value, "xcd_local": bool(xcd.value)}


def measure_rates() -> dict:
    """Measured device rates (FP32 FMA TFLOP/s without packed-FP32 instructions,
    LDS read GB/s, HBM read GB/s) from streaming kernels: the peaks the solvers'
    rooflines are read against."""
    v, l, h = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    check(lib().gbdpcg_measure_rates(ctypes.byref(v), ctypes.byref(l), ctypes.byref(h)))
    return {"valu_f32_tflops": v.value, "lds_read_gbs": l.value, "hbm_read_gbs": h.value}


def lds_bytes(dtype, n: int, knots_per_wg: int, wgs: int) -> int:
    return lib().gbdpcg_lds_bytes(_dtype_code(dtype), n, knots_per_wg, wgs)


# ---------------------------------------------------------------------------
# batched plans
# ---------------------------------------------------------------------------

class Batch:
    """A batched plan over device tensors.  `systems` is a list of dicts with keys
    S, Pinv, gamma, lam, n, N and optionally glam, ggamma, gS, tol, max_iter,
    forward (default True), adjoint (default False), wgs."""

    def __init__(self, systems: Sequence[dict], *, exit_mode=EXIT_REL_RESID, recurrence=RECUR_REFERENCE,
                 lds_budget: int = 0, tol=1e-5, max_iter=173, knots_per_wg: int = 0, threads: int = 0,
                 kernel: int = KERNEL_AUTO, autotune: bool = False, trials: int = 3, stream=None,
                 split_launches: bool = False, exchange_scope: int = EXCHANGE_AUTO):
        if not systems:
            raise ValueError("empty batch")
        self._keep = list(systems)
        dt = _dtype_code(systems[0]["S"].dtype)
        arr = (SystemDesc * len(systems))()
        for i, s in enumerate(systems):
            fwd, adj = s.get("forward", True), s.get("adjoint", False)
            if s.get("Pinv") is None or (fwd and s.get("gamma") is None) or \
                    (adj and (s.get("glam") is None or s.get("ggamma") is None or s.get("gS") is None)):
                raise ValueError(f"system {i}: needs Pinv, gamma (forward) and glam, ggamma, gS (adjoint)")
            if s["S"].dtype != systems[0]["S"].dtype or s["S"].device != systems[0]["S"].device:
                raise ValueError(f"system {i}: every system of a batch shares the first one's dtype and device")
            _check_system(s["S"], s["n"], s["N"], mats=(("Pinv", s["Pinv"]), ("gS", s.get("gS"))),
                          vecs=(("gamma", s.get("gamma")), ("lam", s["lam"]), ("glam", s.get("glam")),
                                ("ggamma", s.get("ggamma"))), where=f"system {i}: ")
            d = arr[i]
            d.S, d.Pinv, d.gamma, d.lam = _dev(s["S"], "S"), _dev(s["Pinv"], "Pinv"), _dev(s.get("gamma"), "gamma"), \
                _dev(s["lam"], "lam")
            d.glambda, d.ggamma, d.gS = _dev(s.get("glam"), "glam"), _dev(s.get("ggamma"), "ggamma"), \
                _dev(s.get("gS"), "gS")
            d.n, d.N = int(s["n"]), int(s["N"])
            d.exit_tol = float(s.get("tol", tol))
            d.max_iter = int(s.get("max_iter", max_iter))
            d.flags = (SYS_FORWARD if s.get("forward", True) else 0) | (SYS_ADJOINT if s.get("adjoint", False) else 0)
            d.wgs = int(s.get("wgs", 0))
        cfg = make_config(exit_mode=exit_mode, recurrence=recurrence, stream=_stream_handle(stream), kernel=kernel,
                          knots_per_wg=knots_per_wg, threads=threads, batch_split=split_launches,
                          exchange_scope=exchange_scope)
        h = ctypes.c_void_p()
        self.tune_report = None
        if autotune:
            rep = TuneReport()
            check(lib().gbdpcg_batch_autotune(ctypes.byref(h), dt, arr, len(systems), ctypes.byref(cfg), int(trials),
                                              ctypes.byref(rep)))
            self.tune_report = rep.as_dict()
        else:
            check(lib().gbdpcg_batch_create(ctypes.byref(h), dt, arr, len(systems), ctypes.byref(cfg),
                                            int(lds_budget)))
        self._h = h
        self.count = len(systems)

    def run(self, stream=None) -> None:
        check(lib().gbdpcg_batch_run(self._h, _stream_handle(stream)))

    def warm_start(self, seed: int, noise_scale: float = 0.0, stream=None) -> None:
        """lambda shifted one knot towards the start, gamma += noise_scale * counter
        normals (seed, system index) for every forward system, in one launch."""
        check(lib().gbdpcg_batch_warm_start(self._h, int(seed), float(noise_scale), _stream_handle(stream)))

    def results(self, stream=None):
        res = (Result * self.count)()
        check(lib().gbdpcg_batch_results(self._h, res, _stream_handle(stream)))
        return [r for r in res]

    def info(self) -> dict:
        bi = BatchInfo()
        check(lib().gbdpcg_batch_info_get(self._h, ctypes.byref(bi)))
        return {f: getattr(bi, f) for f, _ in bi._fields_ if f != "reserved"}

    def wgs(self):
        out = (ctypes.c_uint32 * self.count)()
        check(lib().gbdpcg_batch_wgs(self._h, out))
        return list(out)

    def close(self):
        if getattr(self, "_h", None):
            lib().gbdpcg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
