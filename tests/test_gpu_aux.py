"""GPU tests of the entry points around the solve: on-device preconditioner
formation, the LQR Schur complement kernel, the warm-start shift, the
host-pointer and CSR overloads, the async result record, both kernels forced
at every specialised state size, and the error paths.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _tdt(dtype):
    import torch
    return torch.float64 if dtype == np.float64 else torch.float32


def dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).astype(dtype))).to("cuda", _tdt(dtype))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def lqr(n, N, seed):
    import gbdpcg
    return gbdpcg.lqr_generate(n, max(n // 2, 1), N, seed, dtype=np.float64)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("n", [1, 2, 6, 14, 28])
def test_precond_device_matches_definition(gpu, dtype, kind, n):
    """Block-Jacobi / symmetric-stair Pinv built on the device against the FP64
    definition (dense inverses of the diagonal blocks); pads stay zero."""
    import torch
    import gbdpcg
    N = 33
    S, _ = lqr(n, N, 300 + n)
    ref = O.precond(S, n, N, kind)
    dS = dev(S, dtype)
    dP = torch.full_like(dS, float("nan"))
    gbdpcg.precond_device(dS, dP, n, N, kind)
    P = dP.cpu().numpy()
    mask = O.pattern_mask(n, N).astype(bool)
    assert np.all(P[~mask] == 0)
    if kind == 1:  # Jacobi: the off-diagonal blocks are zero
        off = np.zeros((N, 3, n * n), bool)
        off[:, 0] = off[:, 2] = True
        assert np.all(P[off.reshape(-1)] == 0)
    assert rel(P, ref) <= (1e-4 if dtype == np.float32 else 1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind", [1, 2])
def test_precond_device_every_size(gpu, dtype, kind):
    """Every n from 1 to 40 and n = 48 (the one-wave register inversions of
    n <= 32 at each of their four widths, and the LDS Gauss-Jordan form above
    32), FP32 also n = 64, against the FP64 definition; pads stay zero."""
    import torch
    import gbdpcg
    N = 7
    sizes = list(range(1, 41)) + [48] + ([64] if dtype == np.float32 else [])
    for n in sizes:
        S, _ = lqr(n, N, 700 + n)
        ref = O.precond(S, n, N, kind)
        dS = dev(S, dtype)
        dP = torch.full_like(dS, float("nan"))
        gbdpcg.precond_device(dS, dP, n, N, kind)
        P = dP.cpu().numpy()
        mask = O.pattern_mask(n, N).astype(bool)
        assert np.all(P[~mask] == 0), n
        assert rel(P, ref) <= (1e-5 if dtype == np.float32 else 1e-12), (n, rel(P, ref))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lqr_schur_device_every_size(gpu, dtype):
    """S from (A, B, Q, R) on the device for every n from 1 to 32, with m = n / 2
    inputs (at least one) and with none, against the host construction."""
    import torch
    import gbdpcg
    N = 6
    for n in range(1, 33):
        for m in sorted({max(n // 2, 1), 0}):
            S_h, _, (A, B, Q, R) = gbdpcg.lqr_generate(n, m, N, 900 + n, dtype=np.float64, with_lqr=True)
            dS = torch.full((3 * n * n * N,), float("nan"), dtype=_tdt(dtype), device="cuda")
            gbdpcg.lqr_schur_device(dev(A, dtype), dev(B, dtype) if m else None, dev(Q, dtype),
                                    dev(R, dtype) if m else None, dS, n, m, N)
            got = dS.cpu().numpy()
            assert np.isfinite(got).all(), (n, m)
            assert rel(got, S_h) <= (1e-5 if dtype == np.float32 else 1e-13), (n, m, rel(got, S_h))
            assert np.all(got[~O.pattern_mask(n, N).astype(bool)] == 0), (n, m)


def test_precond_device_singular_block(gpu):
    import gbdpcg
    n, N = 3, 4
    S, _ = lqr(n, N, 5)
    S = S.reshape(N, 3, n * n).copy()
    S[2, 1] = 0.0  # theta_2 = 0
    dS = dev(S.reshape(-1), np.float64)
    dP = dS.clone()
    with pytest.raises(gbdpcg.GbdpcgError) as e:
        gbdpcg.precond_device(dS, dP, n, N, gbdpcg.PRECOND_SS)
    assert e.value.code == gbdpcg.ERR_INVALID


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("precond", [1, 2])
def test_solve_builds_pinv(gpu, dtype, precond):
    """Pinv == NULL: the solver forms cfg.precond on the device; the result matches
    a solve given the oracle's Pinv."""
    import gbdpcg
    n, N = 14, 64
    S, g = lqr(n, N, 41)
    P = O.precond(S, n, N, precond)
    kw = dict(tol=1e-5 if dtype == np.float32 else 1e-10, max_iter=300, exit_mode=gbdpcg.EXIT_REL_RESID,
              recurrence=gbdpcg.RECUR_FUSED, precond=precond)
    lam_a = dev(np.zeros(n * N), dtype)
    ra = gbdpcg.solve(dev(S, dtype), dev(g, dtype), lam_a, n, N, None, **kw)
    lam_b = dev(np.zeros(n * N), dtype)
    rb = gbdpcg.solve(dev(S, dtype), dev(g, dtype), lam_b, n, N, dev(P, dtype), **kw)
    assert ra.status == rb.status == gbdpcg.CONVERGED
    assert abs(ra.iters - rb.iters) <= 1
    exact = O.dense_solve(S, g, n, N)
    lim = 1e-3 if dtype == np.float32 else 1e-8
    assert rel(lam_a.cpu().numpy(), exact) <= lim and rel(lam_b.cpu().numpy(), exact) <= lim


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n,m", [(2, 1), (6, 3), (14, 7), (28, 14), (4, 0)])
def test_lqr_schur_device(gpu, dtype, n, m):
    """S from (A, B, Q, R) on the device against the host construction."""
    import torch
    import gbdpcg
    N = 17
    S_h, _, (A, B, Q, R) = gbdpcg.lqr_generate(n, m, N, 60 + n, dtype=np.float64, with_lqr=True)
    dS = torch.full((3 * n * n * N,), float("nan"), dtype=_tdt(dtype), device="cuda")
    gbdpcg.lqr_schur_device(dev(A, dtype), dev(B, dtype) if m else None, dev(Q, dtype),
                            dev(R, dtype) if m else None, dS, n, m, N)
    got = dS.cpu().numpy()
    assert np.isfinite(got).all()
    assert rel(got, S_h) <= (1e-5 if dtype == np.float32 else 1e-13)
    mask = O.pattern_mask(n, N).astype(bool)
    assert np.all(got[~mask] == 0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n,N", [(14, 128), (28, 512), (3, 1), (5, 2)])
def test_warm_start_device(gpu, dtype, n, N):
    """lambda[k] <- lambda[k+1] (last knot kept), gamma += noise: exact."""
    import gbdpcg
    rng = np.random.default_rng(n * 1000 + N)
    lam = rng.standard_normal(n * N).astype(dtype)
    gam = rng.standard_normal(n * N).astype(dtype)
    noise = (1e-2 * rng.standard_normal(n * N)).astype(dtype)
    dl, dg = dev(lam, dtype), dev(gam, dtype)
    gbdpcg.warm_start_device(dl, dg, n, N, dev(noise, dtype))
    expect = lam.copy()
    expect[: n * (N - 1)] = lam[n:]
    assert np.array_equal(dl.cpu().numpy(), expect)
    assert np.array_equal(dg.cpu().numpy(), gam + noise)
    gbdpcg.warm_start_device(dl, dg, n, N, None)  # no noise: gamma unchanged
    assert np.array_equal(dg.cpu().numpy(), gam + noise)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_host_pointer_solve_matches_device(gpu, dtype):
    """The reference's host-pointer call returns the iteration count and the same
    lambda, bit for bit, as the device-pointer call (same kernel, same plan)."""
    import gbdpcg
    n, N = 14, 128
    S, g = lqr(n, N, 77)
    P = O.precond(S, n, N, 2).astype(dtype)
    S, g = S.astype(dtype), g.astype(dtype)
    kw = dict(tol=1e-5, max_iter=173, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    lam_h = np.zeros(n * N, dtype)
    it = gbdpcg.solve_host(S, g, lam_h, n, N, P, **kw)
    lam_d = dev(np.zeros(n * N), dtype)
    r = gbdpcg.solve(dev(S, dtype), dev(g, dtype), lam_d, n, N, dev(P, dtype), **kw)
    assert it == r.iters and np.array_equal(lam_h, lam_d.cpu().numpy())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_csr_solve(gpu, dtype):
    """CSR overload (stubbed in the reference): same result as the cbtd call."""
    sp = pytest.importorskip("scipy.sparse")
    import gbdpcg
    n, N = 6, 20
    S, g = lqr(n, N, 88)
    P = O.precond(S, n, N, 2)
    Sc = sp.csr_matrix(O.cbtd_to_dense(S, n, N))
    Pc = sp.csr_matrix(O.cbtd_to_dense(P, n, N))
    kw = dict(tol=1e-9 if dtype == np.float64 else 1e-5, max_iter=200, exit_mode=gbdpcg.EXIT_REL_RESID)
    lam_c = np.zeros(n * N, dtype)
    it_c = gbdpcg.solve_csr((Sc.indptr, Sc.indices, Sc.data), g, lam_c, n, N, (Pc.indptr, Pc.indices, Pc.data),
                            **kw)
    lam_b = np.zeros(n * N, dtype)
    it_b = gbdpcg.solve_host(S.astype(dtype), g.astype(dtype), lam_b, n, N, P.astype(dtype), **kw)
    assert it_c == it_b and np.array_equal(lam_c, lam_b)
    # no Pinv: built from S on the device
    lam_n = np.zeros(n * N, dtype)
    it_n = gbdpcg.solve_csr((Sc.indptr, Sc.indices, Sc.data), g, lam_n, n, N, None, **kw)
    assert abs(it_n - it_b) <= 1
    assert rel(lam_n, O.dense_solve(S, g, n, N)) <= (1e-3 if dtype == np.float32 else 1e-7)


def test_async_result_record(gpu):
    """solve_device_async writes the same result record to device memory as the
    synchronous call returns."""
    import torch
    import gbdpcg
    n, N, dtype = 14, 96, np.float32
    S, g = lqr(n, N, 13)
    P = O.precond(S, n, N, 2)
    kw = dict(tol=1e-5, max_iter=173, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    dS, dP, dg = dev(S, dtype), dev(P, dtype), dev(g, dtype)
    l1 = dev(np.zeros(n * N), dtype)
    r1 = gbdpcg.solve(dS, dg, l1, n, N, dP, **kw)
    l2 = dev(np.zeros(n * N), dtype)
    rec = torch.zeros(gbdpcg.RESULT_BYTES, dtype=torch.uint8, device="cuda")
    gbdpcg.solve_async(dS, dg, l2, n, N, dP, rec, **kw)
    torch.cuda.synchronize()
    r2 = gbdpcg.Result.from_buffer_copy(rec.cpu().numpy().tobytes())
    assert (r2.iters, r2.status) == (r1.iters, r1.status) and r2.rr == r1.rr and r2.eta == r1.eta
    assert torch.equal(l1, l2)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [2, 6, 12, 14, 28])
@pytest.mark.parametrize("kernel", [1, 2, 3])
@pytest.mark.parametrize("recurrence", [0, 1])
def test_forced_kernels(gpu, dtype, n, kernel, recurrence):
    """Both recurrences on the LDS-resident kernel and on both register-resident
    layouts (column chunks; row broadcast), each against the oracle (iterations
    within 1, lambda to the precision bar)."""
    import gbdpcg
    N = 64
    S, g = lqr(n, N, 500 + n)
    P = O.precond(S, n, N, 2)
    Sd, Pd, gd = S.astype(dtype), P.astype(dtype), g.astype(dtype)
    tol = 1e-5 if dtype == np.float32 else 1e-10
    lam = dev(np.zeros(n * N), dtype)
    try:
        r = gbdpcg.solve(dev(Sd, dtype), dev(gd, dtype), lam, n, N, dev(Pd, dtype), tol=tol, max_iter=400,
                         exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=recurrence, kernel=kernel)
    except gbdpcg.GbdpcgError as e:
        # the register kernel holds FP64 rows of n > 14 only at 256 threads, and the
        # row-broadcast layout takes n in {12, 14} only; they may decline, but only
        # with UNSUPPORTED
        assert e.code == gbdpcg.ERR_UNSUPPORTED
        assert (kernel == 2 and dtype == np.float64 and n > 14) or (kernel == 3 and n not in (12, 14))
        return
    assert kernel != 3 or n in (12, 14)
    assert r.status == gbdpcg.CONVERGED
    _, it_same, *_ = O.pcg(Sd, Pd, gd, np.zeros(n * N, dtype), n, N, tol, 400, O.EXIT_REL, dtype,
                           fused=recurrence == 1)
    assert abs(r.iters - it_same) <= 1
    lam_ref, *_ = O.pcg(S, P, g, np.zeros(n * N), n, N, 0.0, r.iters)
    assert rel(lam.cpu().numpy(), lam_ref) <= (1e-4 if dtype == np.float32 else 1e-10)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [12, 14])
@pytest.mark.parametrize("N", [4, 10, 64, 128, 256])
@pytest.mark.parametrize("recurrence", [0, 1])
def test_one_wave_ranks(gpu, dtype, n, N, recurrence):
    """Row-broadcast kernel at 64 threads: one wave per rank, two knots each (128
    ranks at N = 256: lanes poll a second rank each after the first), against the
    oracle like every forced layout.  Odd N does not split into ranks of two
    knots: the request is declined."""
    import gbdpcg
    S9, g9 = lqr(n, 9, 9)
    with pytest.raises(gbdpcg.GbdpcgError) as e:
        gbdpcg.solve(dev(S9, dtype), dev(g9, dtype), dev(np.zeros(n * 9), dtype), n, 9, None, recurrence=recurrence,
                     kernel=gbdpcg.KERNEL_ROWCAST, threads=64)
    assert e.value.code == gbdpcg.ERR_UNSUPPORTED
    S, g = lqr(n, N, 900 + n + N)
    P = O.precond(S, n, N, 2)
    Sd, Pd, gd = S.astype(dtype), P.astype(dtype), g.astype(dtype)
    tol = 1e-5 if dtype == np.float32 else 1e-10
    lam = dev(np.zeros(n * N), dtype)
    r = gbdpcg.solve(dev(Sd, dtype), dev(gd, dtype), lam, n, N, dev(Pd, dtype), tol=tol, max_iter=400,
                     exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=recurrence, kernel=gbdpcg.KERNEL_ROWCAST, threads=64)
    assert r.status == gbdpcg.CONVERGED
    _, it_same, *_ = O.pcg(Sd, Pd, gd, np.zeros(n * N, dtype), n, N, tol, 400, O.EXIT_REL, dtype,
                           fused=recurrence == 1)
    assert abs(r.iters - it_same) <= 1
    lam_ref, *_ = O.pcg(S, P, g, np.zeros(n * N), n, N, 0.0, r.iters)
    assert rel(lam.cpu().numpy(), lam_ref) <= (1e-4 if dtype == np.float32 else 1e-10)


@pytest.mark.parametrize("wgs,threads", [(1, 256), (16, 256), (16, 512), (64, 64), (200, 256)])
def test_exchange_floor(gpu, wgs, threads):
    """The exchange-only rounds finish (every wait ends) and give a plausible
    per-round latency; bad arguments are rejected."""
    import gbdpcg
    r = gbdpcg.exchange_floor(wgs, threads, rounds=500)
    print(f"exchange floor: {wgs} x {threads}: {r['us_per_round']:.3f} us/round, xcd_local={r['xcd_local']}")
    assert 0.0 < r["us_per_round"] < 20.0
    assert not (wgs == 1 and r["xcd_local"])
    for bad in (dict(wgs=0, threads=256), dict(wgs=4, threads=128), dict(wgs=4, threads=256, rounds=5)):
        with pytest.raises(gbdpcg.GbdpcgError) as e:
            gbdpcg.exchange_floor(**bad)
        assert e.value.code == gbdpcg.ERR_INVALID


def test_register_kernel_declines_unsupported(gpu):
    import gbdpcg
    n, N = 7, 16
    S, g = lqr(n, N, 3)
    lam = dev(np.zeros(n * N), np.float32)
    with pytest.raises(gbdpcg.GbdpcgError) as e:
        gbdpcg.solve(dev(S, np.float32), dev(g, np.float32), lam, n, N, None, recurrence=gbdpcg.RECUR_FUSED,
                     kernel=gbdpcg.KERNEL_REGISTER)
    assert e.value.code == gbdpcg.ERR_UNSUPPORTED


@pytest.mark.parametrize("recurrence,kernel", [(0, 0), (1, 1), (1, 2), (0, 3), (1, 3)])
def test_nonfinite_input(gpu, recurrence, kernel):
    """A NaN in gamma ends the solve with status NONFINITE (no hang, no exception)."""
    import gbdpcg
    n, N = 14, 32
    S, g = lqr(n, N, 21)
    g[5] = np.nan
    lam = dev(np.zeros(n * N), np.float32)
    r = gbdpcg.solve(dev(S, np.float32), dev(g, np.float32), lam, n, N, None, tol=1e-5, max_iter=50,
                     exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=recurrence, kernel=kernel)
    assert r.status == gbdpcg.NONFINITE


def test_check_occupancy(gpu):
    """gbdpcg_check_occupancy (the reference's checkPcgOccupancy,
    include/pcg.cuh:23-49, which exits instead of returning): OK with the
    co-resident workgroup capacity (a multiple of the CU count) for a system
    that fits, TOO_LARGE for one whose block row does not fit LDS even at one
    knot per workgroup, INVALID for bad arguments."""
    import ctypes
    import torch
    import gbdpcg
    L = gbdpcg.lib()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    cap = ctypes.c_uint32(0)
    assert L.gbdpcg_check_occupancy(gbdpcg.F32, 14, 128, ctypes.byref(cap)) == 0
    assert cap.value >= cus and cap.value % cus == 0
    for n, N in ((2, 3), (28, 512)):
        assert L.gbdpcg_check_occupancy(gbdpcg.F64, n, N, ctypes.byref(cap)) == 0 and cap.value % cus == 0
    # n = 64 FP64: one knot's S and Phi^-1 rows alone exceed the 160 KiB LDS
    assert gbdpcg.lds_bytes(np.float64, 64, 1, 2) > 160 * 1024
    assert L.gbdpcg_check_occupancy(gbdpcg.F64, 64, 8, ctypes.byref(cap)) == gbdpcg.ERR_TOO_LARGE
    assert L.gbdpcg_check_occupancy(gbdpcg.F32, 0, 8, ctypes.byref(cap)) == gbdpcg.ERR_INVALID
    assert L.gbdpcg_check_occupancy(7, 14, 8, ctypes.byref(cap)) == gbdpcg.ERR_INVALID


def test_invalid_arguments(gpu):
    import gbdpcg
    n, N = 4, 8
    S, g = lqr(n, N, 1)
    dS, dg, lam = dev(S, np.float32), dev(g, np.float32), dev(np.zeros(n * N), np.float32)
    import ctypes
    for kw in (dict(n=0, N=N), dict(n=n, N=0)):
        with pytest.raises(ValueError):  # the binding checks sizes before calling the library
            gbdpcg.solve(dS, dg, lam, kw["n"], kw["N"], None)
        # and the C ABI rejects them itself
        cfg = gbdpcg.make_config()
        rc = gbdpcg.lib().gbdpcg_solve_device_f32(kw["n"], kw["N"], dS.data_ptr(), None, dg.data_ptr(),
                                                  lam.data_ptr(), ctypes.byref(cfg), None)
        assert rc == gbdpcg.ERR_INVALID
    with pytest.raises(gbdpcg.GbdpcgError) as e:
        gbdpcg.solve(dS, dg, lam, n, N, None, max_iter=10 ** 6)
    assert e.value.code == gbdpcg.ERR_INVALID
    with pytest.raises(gbdpcg.GbdpcgError) as e:
        gbdpcg.solve(dS, dg, lam, n, N, None, exit_mode=7)
    assert e.value.code == gbdpcg.ERR_INVALID
    # a failed call leaves lambda untouched and the next call works
    assert not lam.cpu().numpy().any()
    r = gbdpcg.solve(dS, dg, lam, n, N, None, tol=1e-5, max_iter=100, exit_mode=gbdpcg.EXIT_REL_RESID)
    assert r.status == gbdpcg.CONVERGED


def test_measure_rates(gpu):
    """The resource-rate kernels finish and give rates in the device's range
    (nominal: FP32 FMA 78.6 TFLOP/s without packed-FP32, LDS reads 79-157 TB/s,
    HBM 8 TB/s)."""
    import gbdpcg
    r = gbdpcg.measure_rates()
    print(f"rates: {r}")
    assert 10.0 < r["valu_f32_tflops"] < 120.0
    assert 5_000.0 < r["lds_read_gbs"] < 250_000.0
    assert 1_000.0 < r["hbm_read_gbs"] < 10_000.0
