"""Host-side checks of the product library (no GPU needed).

* the C-ABI library loads and exports every entry point include/gbdpcg.h declares;
* the LQR generator: deterministic, symmetric positive definite, and equal to the
  Schur complement C G^-1 C^T of the LQR KKT system built densely here from the
  generator's own A, B, Q, R;
* the host preconditioner matches the oracle's restatement of the definition;
* CSR -> cbtd conversion, including rejection of entries outside the pattern;
* argument validation returns error codes (and never needs the device);
* the ctypes mirrors of the C structs have the header's sizes and field offsets.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import gbdpcg
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "gbdpcg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gbdpcg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = gbdpcg.lib()
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    bound = {n for n, _, _ in gbdpcg.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_version_and_defaults():
    assert gbdpcg.version().startswith("gbdpcg")
    cfg = gbdpcg.Config()
    gbdpcg.lib().gbdpcg_config_init(ctypes.byref(cfg))
    # reference defaults, include/constants.cuh:14-20 and include/types.cuh:33
    assert cfg.exit_tol == 1e-6 and cfg.max_iter == 25 and cfg.block[0] == 64 and cfg.empty_pinv == 1
    assert cfg.exit_mode == gbdpcg.EXIT_ABS_ETA and cfg.recurrence == gbdpcg.RECUR_REFERENCE
    assert ctypes.sizeof(gbdpcg.Result) == 48


def test_generator_is_deterministic():
    a = gbdpcg.lqr_generate(6, 3, 9, 1234, dtype=np.float64)
    b = gbdpcg.lqr_generate(6, 3, 9, 1234, dtype=np.float64)
    c = gbdpcg.lqr_generate(6, 3, 9, 1235, dtype=np.float64)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    f = gbdpcg.lqr_generate(6, 3, 9, 1234, dtype=np.float32)
    assert np.array_equal(f[0], a[0].astype(np.float32))


@pytest.mark.parametrize("n,m,N", [(2, 1, 3), (6, 3, 7), (14, 7, 5)])
def test_generator_is_kkt_schur_complement(n, m, N):
    S, g, (A, B, Q, R) = gbdpcg.lqr_generate(n, m, N, 77 + n, dtype=np.float64, with_lqr=True)
    # KKT of min sum x'Q x + u'R u s.t. x_0 = .., x_k = A_{k-1} x_{k-1} + B_{k-1} u_{k-1}
    # variables z = [x_0, u_0, x_1, u_1, ..., x_{N-1}] ; S = C G^-1 C^T
    nz = N * n + (N - 1) * m
    off = lambda k: k * (n + m)  # x_k offset; u_k at off(k) + n
    C = np.zeros((N * n, nz))
    Gd = np.zeros(nz)
    for k in range(N):
        C[k * n:(k + 1) * n, off(k):off(k) + n] = np.eye(n)
        Gd[off(k):off(k) + n] = Q[k]
        if k > 0:
            C[k * n:(k + 1) * n, off(k - 1):off(k - 1) + n] = -A[k - 1]
            C[k * n:(k + 1) * n, off(k - 1) + n:off(k - 1) + n + m] = -B[k - 1]
        if k < N - 1:
            Gd[off(k) + n:off(k) + n + m] = R[k]
    Sd = C @ np.diag(1 / Gd) @ C.T
    assert np.allclose(O.cbtd_to_dense(S, n, N), Sd, rtol=1e-12, atol=1e-12)
    assert np.all(np.linalg.eigvalsh(Sd) > 0)
    # drawn ranges
    assert np.all((Q >= 0.1) & (Q <= 10)) and np.all((R >= 0.01) & (R <= 1))
    assert abs(np.mean(A.reshape(N, -1) - np.eye(n).reshape(1, -1))) < 0.05


def test_randn_moments():
    x = gbdpcg.randn(200000, 5, dtype=np.float64)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01


@pytest.mark.parametrize("kind", [gbdpcg.PRECOND_JACOBI, gbdpcg.PRECOND_SS])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_host_preconditioner_matches_definition(kind, dtype):
    n, N = 6, 8
    S, _ = gbdpcg.lqr_generate(n, 3, N, 4, dtype=np.float64)
    P = gbdpcg.precond_host(S.astype(dtype), n, N, kind)
    ref = O.precond(S.astype(dtype).astype(np.float64), n, N, kind)
    tol = 1e-10 if dtype == np.float64 else 2e-5
    assert np.allclose(P, ref, rtol=tol, atol=tol * np.abs(ref).max())
    pads = P.reshape(N, 3, n * n)
    assert not pads[0, 0].any() and not pads[-1, 2].any()


def test_host_preconditioner_singular_block():
    S = np.zeros(3 * 4 * 2, dtype=np.float64)
    with pytest.raises(gbdpcg.GbdpcgError) as e:
        gbdpcg.precond_host(S, 2, 2, gbdpcg.PRECOND_SS)
    assert e.value.code == gbdpcg.ERR_INVALID


def test_csr_to_cbtd():
    n, N = 3, 5
    S, _ = gbdpcg.lqr_generate(n, 1, N, 9, dtype=np.float64)
    D = O.cbtd_to_dense(S, n, N)
    M = sp.csr_matrix(D)
    C = gbdpcg.csr_to_cbtd(M.indptr, M.indices, M.data, n, N)
    assert np.array_equal(C, S)
    # duplicates are summed (CSR convention)
    rows = np.array([0, 2, 2, 2, 2, 2, 2])
    Dd = np.zeros((2, 2))
    Dd[0, 0] = 1.0
    Dd[1, 1] = 5.0
    indptr = np.array([0, 1, 3])
    indices = np.array([0, 1, 1])
    data = np.array([1.0, 2.0, 3.0])
    C2 = gbdpcg.csr_to_cbtd(indptr, indices, data, 1, 2)
    assert np.allclose(O.cbtd_to_dense(C2, 1, 2), Dd)
    # an entry two blocks off the diagonal is outside the pattern
    bad = D.copy()
    bad[0, 3 * n - 1] = 1.0
    Mb = sp.csr_matrix(bad)
    with pytest.raises(gbdpcg.GbdpcgError) as e:
        gbdpcg.csr_to_cbtd(Mb.indptr, Mb.indices, Mb.data, n, N)
    assert e.value.code == gbdpcg.ERR_UNSUPPORTED
    del rows


def test_lds_bytes_formula():
    # grows with knots per workgroup and with n; FP64 is twice FP32
    a = gbdpcg.lds_bytes(np.float32, 14, 8, 16)
    assert a > 0 and gbdpcg.lds_bytes(np.float64, 14, 8, 16) == 2 * a
    assert gbdpcg.lds_bytes(np.float32, 14, 9, 16) > a
    assert gbdpcg.lds_bytes(np.float32, 0, 8, 16) == 0


def test_argument_validation_without_device():
    L = gbdpcg.lib()
    cfg = gbdpcg.make_config()
    res = gbdpcg.Result()
    # null device pointers
    assert L.gbdpcg_solve_device_f32(14, 128, None, None, None, None, ctypes.byref(cfg), ctypes.byref(res)) == \
        gbdpcg.ERR_INVALID
    # zero sizes
    buf = (ctypes.c_float * 8)()
    assert L.gbdpcg_solve_f32(buf, None, buf, buf, 0, 3, ctypes.byref(cfg)) == gbdpcg.ERR_INVALID
    # bad enums and limits
    for field, val in (("exit_mode", 7), ("recurrence", 9), ("precond", 0), ("max_iter", 10 ** 6)):
        c2 = gbdpcg.make_config()
        setattr(c2, field, val)
        assert L.gbdpcg_solve_f32(buf, None, buf, buf, 2, 3, ctypes.byref(c2)) == gbdpcg.ERR_INVALID
        assert gbdpcg.last_error()
    # null config
    assert L.gbdpcg_solve_f32(buf, None, buf, buf, 2, 3, None) == gbdpcg.ERR_INVALID
    # batch: empty and invalid descriptors
    h = ctypes.c_void_p()
    assert L.gbdpcg_batch_create(ctypes.byref(h), gbdpcg.F32, None, 0, ctypes.byref(cfg), 0) == gbdpcg.ERR_INVALID
    d = (gbdpcg.SystemDesc * 1)()
    assert L.gbdpcg_batch_create(ctypes.byref(h), gbdpcg.F32, d, 1, ctypes.byref(cfg), 0) == gbdpcg.ERR_INVALID
    assert L.gbdpcg_batch_destroy(None) == 0


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes Structure of the binding against a C program compiled with the
    header: size and the offset of every field."""
    structs = {"gbdpcg_config": gbdpcg.Config, "gbdpcg_result": gbdpcg.Result,
               "gbdpcg_system_desc": gbdpcg.SystemDesc, "gbdpcg_batch_info": gbdpcg.BatchInfo,
               "gbdpcg_tune_report": gbdpcg.TuneReport}
    c_name = {"lam": "lambda"}  # Python keyword in the mirror
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "gbdpcg.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {c_name.get(field, field)}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, key, val = line.split()
        got[(cname, key)] = int(val)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)


def test_binding_rejects_mismatched_tensors_before_launch():
    """The binding checks dtype, element count and device of every tensor before
    any launch (a short or mistyped buffer would otherwise be read or written out
    of bounds on the device).  CPU tensors suffice: the checks run first."""
    import torch
    n, N = 6, 8
    S = torch.zeros(3 * n * n * N, dtype=torch.float64)
    vec = torch.zeros(n * N, dtype=torch.float64)
    cases = [
        dict(gamma=vec.float(), lam=vec),                        # gamma in FP32 with FP64 S
        dict(gamma=vec, lam=vec[:-1].clone()),                   # short lam
        dict(gamma=torch.zeros(n * N + 1, dtype=torch.float64), lam=vec),
        dict(gamma=vec, lam=vec, Pinv=S[:-n].clone()),           # short Pinv
    ]
    for kw in cases:
        with pytest.raises(ValueError):
            gbdpcg.solve(S, kw["gamma"], kw["lam"], n, N, kw.get("Pinv"))
        with pytest.raises(ValueError):
            gbdpcg.solve_async(S, kw["gamma"], kw["lam"], n, N, kw.get("Pinv"))
    with pytest.raises(ValueError):  # n N mismatch against S
        gbdpcg.solve(S, vec, vec, n, N + 1)
    with pytest.raises(ValueError):  # adjoint outputs: short gS, FP32 ggamma
        gbdpcg.adjoint(S, vec, vec, vec, vec, S[:-1].clone(), n, N, S)
    with pytest.raises(ValueError):
        gbdpcg.adjoint(S, vec, vec, vec, vec.float(), S, n, N, S)
    with pytest.raises(ValueError):  # batch descriptors are checked per system
        gbdpcg.Batch([dict(S=S, Pinv=S, gamma=vec, lam=vec[:-1].clone(), n=n, N=N)])
    with pytest.raises(ValueError):
        gbdpcg.Batch([dict(S=S, Pinv=S, gamma=vec, lam=vec, n=n, N=N, adjoint=True, glam=vec,
                           ggamma=vec, gS=None)])
    # host-pointer call: sizes against (n, N) before anything is copied
    Sh = np.zeros(3 * n * n * N)
    with pytest.raises(ValueError):
        gbdpcg.solve_host(Sh, np.zeros(n * N), np.zeros(n * N - 1), n, N)
    with pytest.raises(ValueError):
        gbdpcg.solve_host(Sh[:-1].copy(), np.zeros(n * N), np.zeros(n * N), n, N)


def test_exchange_scope_is_validated():
    L = gbdpcg.lib()
    buf = (ctypes.c_float * 8)()
    cfg = gbdpcg.make_config(exchange_scope=2)
    assert L.gbdpcg_solve_f32(buf, None, buf, buf, 2, 3, ctypes.byref(cfg)) == gbdpcg.ERR_INVALID
    assert "exchange_scope" in gbdpcg.last_error()
    h = ctypes.c_void_p()
    d = (gbdpcg.SystemDesc * 1)()
    assert L.gbdpcg_batch_create(ctypes.byref(h), gbdpcg.F32, d, 1, ctypes.byref(cfg), 0) == gbdpcg.ERR_INVALID
    # the default is the automatic choice
    c0 = gbdpcg.Config()
    L.gbdpcg_config_init(ctypes.byref(c0))
    assert c0.exchange_scope == gbdpcg.EXCHANGE_AUTO
