"""GPU parity under the reference's own exit rule and defaults, its LQR compile
example shape, the block-Jacobi preconditioner, and the device-scope exchange.

* Exit test |r'z| < tol (include/pcg.cuh:195) at n=14, N in {32, 64, 128, 256},
  both recurrences, FP32 and FP64, register- and LDS-resident kernels: iteration count within +-1 of the oracle at the
  same precision (and, for the fused form, of the reference form), lambda within
  1e-4 / 1e-10 of the FP64 oracle run for the same number of iterations.
* The reference defaults, tol 1e-6 and max_iter 25 (include/constants.cuh:14-16):
  where the oracle ends at max_iter well short of the tolerance, the device solve
  ends with status MAX_ITER after exactly 25 iterations (include/pcg.cuh:212).
* LQR n=2, m=1, N=3 with the block-Jacobi Pinv in FP32 (the shape of the
  reference's compile example, examples/pcg_solve.cu:14-41), relative residual
  1e-5, max_iter 173: the device-pointer and the host-pointer calls.
* Block-Jacobi Pinv at n=14, N=128.
* GBDPCG_EXCHANGE_DEVICE (device-scope exchange stores) on the n=14 set and on a
  batch: bit-identical to the automatic (XCD-local) exchange.
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import LAM_TOL, check_against_oracle, gpu_solve, lqr_system, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("N", [32, 64, 128, 256])
@pytest.mark.parametrize("recurrence", [0, 1])
@pytest.mark.parametrize("kernel", [0, 1])
def test_abs_eta_rule_n14(gpu, dtype, N, recurrence, kernel):
    """kernel 0: the automatic choice (register-resident rows); 1: the
    LDS-resident kernel."""
    import gbdpcg
    n = 14
    S, P, g = lqr_system(n, N, 2000 + N, dtype)
    res, lam, it_same = check_against_oracle(S, P, g, n, N, dtype, 1e-5, 173, gbdpcg.EXIT_ABS_ETA, recurrence,
                                             kernel=kernel)
    assert res.status == gbdpcg.CONVERGED
    # the reference rule: the final r'z is below the tolerance in absolute value
    assert abs(res.eta) < 1e-5
    print(f"abs rule n14 N={N} {dtype.__name__} rec={recurrence}: iters gpu={res.iters} oracle={it_same}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("N", [32, 128, 256])
@pytest.mark.parametrize("recurrence", [0, 1])
def test_reference_defaults(gpu, dtype, N, recurrence):
    """gbdpcg.solve with nothing but the recurrence set runs the reference's
    config defaults: tol 1e-6 on |r'z|, max_iter 25."""
    import gbdpcg
    n = 14
    S, P, g = lqr_system(n, N, 2000 + N, dtype)
    lam0 = np.zeros(n * N, dtype)
    lam, res = gpu_solve(S, P, g, lam0, n, N, dtype, recurrence=recurrence)
    fused = recurrence == gbdpcg.RECUR_FUSED
    _, it_same, st_same, eta_same, _, _ = O.pcg(S, P, g, lam0, n, N, 1e-6, 25, O.EXIT_ABS, dtype, fused=fused)
    assert abs(res.iters - it_same) <= 1, (res.iters, it_same)
    if st_same == O.MAX_ITER and abs(eta_same) > 10 * 1e-6:
        # far from the tolerance at max_iter: no rounding difference can end it earlier
        assert (res.status, res.iters) == (gbdpcg.MAX_ITER, 25), (res.status, res.iters)
    lam_ref, *_ = O.pcg(S.astype(np.float64), P.astype(np.float64), g.astype(np.float64), np.zeros(n * N), n, N,
                        0.0, res.iters)
    assert rel(lam, lam_ref) <= LAM_TOL[dtype]
    print(f"defaults n14 N={N} {dtype.__name__} rec={recurrence}: status {res.status} iters {res.iters} "
          f"(oracle {it_same}, status {st_same}, eta {eta_same:.2e})")


@pytest.mark.parametrize("recurrence", [0, 1])
def test_lqr_n2_N3_jacobi_fp32(gpu, recurrence):
    """LQR n=2, m=1, N=3, block-Jacobi Pinv, FP32, relative residual 1e-5,
    max_iter 173: device pointers, host pointers (the reference's call shape,
    include/interface.cuh:24-91) and Pinv formed on the device."""
    import gbdpcg
    n, N, dtype = 2, 3, np.float32
    S, P, g = lqr_system(n, N, 1000, dtype, precond=gbdpcg.PRECOND_JACOBI)
    res, lam, it_same = check_against_oracle(S, P, g, n, N, dtype, 1e-5, 173, gbdpcg.EXIT_REL_RESID, recurrence)
    assert res.status == gbdpcg.CONVERGED
    exact = O.dense_solve(S.astype(np.float64), g.astype(np.float64), n, N)
    assert rel(lam, exact) <= 1e-4
    kw = dict(tol=1e-5, max_iter=173, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=recurrence)
    # host pointers: the same kernel and layout, so the same bits
    lam_h = np.zeros(n * N, dtype)
    it_h = gbdpcg.solve_host(S, g, lam_h, n, N, P, **kw)
    assert it_h == res.iters and np.array_equal(lam_h, lam)
    # no Pinv: block-Jacobi formed on the device from S
    lam_d = np.zeros(n * N, dtype)
    it_d = gbdpcg.solve_host(S, g, lam_d, n, N, None, precond=gbdpcg.PRECOND_JACOBI, **kw)
    assert abs(it_d - it_same) <= 1 and rel(lam_d, exact) <= 1e-4
    print(f"n2 N3 jacobi fp32 rec={recurrence}: iters gpu={res.iters} oracle={it_same} err={rel(lam, exact):.1e}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("recurrence", [0, 1])
def test_jacobi_n14_N128(gpu, dtype, recurrence):
    import gbdpcg
    n, N = 14, 128
    S, P, g = lqr_system(n, N, 2128, dtype, precond=gbdpcg.PRECOND_JACOBI)
    res, lam, it_same = check_against_oracle(S, P, g, n, N, dtype, 1e-5, 173, gbdpcg.EXIT_REL_RESID, recurrence)
    assert res.status == gbdpcg.CONVERGED
    assert rel(lam, O.block_solve(S.astype(np.float64), g.astype(np.float64), n, N)) <= 5e-5
    print(f"jacobi n14 N128 {dtype.__name__} rec={recurrence}: iters gpu={res.iters} oracle={it_same}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("N", [32, 64, 128, 256])
@pytest.mark.parametrize("recurrence", [0, 1])
def test_device_scope_exchange_n14(gpu, dtype, N, recurrence):
    """Device-scope exchange stores (no reliance on a shared XCD L2): parity with
    the oracle, and the same bits as the automatic exchange (the exchange moves
    values, it does not change the arithmetic)."""
    import gbdpcg
    n = 14
    S, P, g = lqr_system(n, N, 2000 + N, dtype)
    res, lam, _ = check_against_oracle(S, P, g, n, N, dtype, 1e-5, 173, gbdpcg.EXIT_REL_RESID, recurrence,
                                       exchange_scope=gbdpcg.EXCHANGE_DEVICE)
    lam_auto, res_auto = gpu_solve(S, P, g, np.zeros(n * N, dtype), n, N, dtype, tol=1e-5, max_iter=173,
                                   exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=recurrence)
    assert res.iters == res_auto.iters and np.array_equal(lam, lam_auto)


@pytest.mark.parametrize("adjoint", [False, True])
def test_device_scope_exchange_batch(gpu, adjoint):
    import gbdpcg
    from test_gpu_batched import check, make_batch
    rng = np.random.default_rng(17)
    specs = [(int(rng.choice([6, 12, 14, 28])), int(rng.integers(16, 97)), 1e-5) for _ in range(16)]
    dtype = np.float32
    systems, host = make_batch(specs, dtype, adjoint=adjoint, seed0=7100)
    kw = dict(exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    b = gbdpcg.Batch(systems, exchange_scope=gbdpcg.EXCHANGE_DEVICE, **kw)
    b.run()
    res = b.results()
    check(systems, host, res, dtype, 1, adjoint=adjoint)
    out_dev = [d["lam"].cpu().numpy().copy() for d in systems]
    for d in systems:
        d["lam"].zero_()
    b2 = gbdpcg.Batch(systems, **kw)
    b2.run()
    assert [r.iters for r in b2.results()] == [r.iters for r in res]
    for d, o in zip(systems, out_dev):
        assert np.array_equal(d["lam"].cpu().numpy(), o)
