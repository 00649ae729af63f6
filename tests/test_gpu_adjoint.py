"""GPU adjoint (differentiation) path against the oracle, FP64 finite differences
and the dense FP64 gradient.

Bars: g_gamma within 1e-4 (FP32) / 1e-10 (FP64) of the FP64 oracle PCG run for the
same number of adjoint iterations; g_S equal (to rounding) to the formula applied
to the returned lambda and g_gamma; with a tight FP64 tolerance, g_S and g_gamma
within 1e-8 of the dense FP64 gradient and directional derivatives within 1e-6 of
central finite differences (eps = 1e-6) on 32 symmetric block-tridiagonal
directions.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _t(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", torch.float64 if dtype == np.float64 else torch.float32)


def system(n, N, seed, dtype):
    import gbdpcg
    S, g = gbdpcg.lqr_generate(n, max(n // 2, 1), N, seed, dtype=np.float64)
    P = O.precond(S, n, N, 2)
    glam = gbdpcg.randn(n * N, seed + 99, dtype=np.float64)
    return S, P, g, glam


def run_adjoint(S, P, g, glam, n, N, dtype, forward=True, lam_in=None, **kw):
    import gbdpcg
    dS, dP, dg, dgl = _t(S.astype(dtype), dtype), _t(P.astype(dtype), dtype), _t(g.astype(dtype), dtype), \
        _t(glam.astype(dtype), dtype)
    lam0 = np.zeros(n * N) if lam_in is None else lam_in
    dl = _t(lam0.astype(dtype), dtype)
    dgg = _t(np.zeros(n * N, dtype), dtype)
    dgS = _t(np.full(3 * n * n * N, np.nan, dtype), dtype)
    res = gbdpcg.adjoint(dS, dg, dl, dgl, dgg, dgS, n, N, dP, forward=forward, **kw)
    return dl.cpu().numpy().astype(np.float64), dgg.cpu().numpy().astype(np.float64), \
        dgS.cpu().numpy().astype(np.float64), res


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("recurrence", [0, 1])
@pytest.mark.parametrize("N", [32, 128])
@pytest.mark.parametrize("mode", [1, 0])
def test_adjoint_against_oracle(gpu, dtype, recurrence, N, mode):
    """mode 1: relative residual; mode 0: the reference's |r'z| < tol
    (include/pcg.cuh:195)."""
    n = 14
    S, P, g, glam = system(n, N, 4000 + N, dtype)
    lam, gg, gS, res = run_adjoint(S, P, g, glam, n, N, dtype, tol=1e-5, max_iter=173, exit_mode=mode,
                                   recurrence=recurrence)
    assert res.status == 0 and res.adj_status == 0
    tol = 1e-4 if dtype == np.float32 else 1e-10
    lam_ref, it, *_ = O.pcg(S, P, g, np.zeros(n * N), n, N, 0.0, res.iters)
    assert rel(lam, lam_ref) <= tol
    gg_ref, it2, *_ = O.pcg(S, P, glam, np.zeros(n * N), n, N, 0.0, res.adj_iters)
    assert rel(gg, gg_ref) <= tol
    # the iteration counts of both solves are within 1 of the same-precision oracle
    _, it_f, *_ = O.pcg(S.astype(dtype), P.astype(dtype), g.astype(dtype), np.zeros(n * N, dtype), n, N, 1e-5, 173,
                        mode, dtype, fused=recurrence == 1)
    _, it_a, *_ = O.pcg(S.astype(dtype), P.astype(dtype), glam.astype(dtype), np.zeros(n * N, dtype), n, N, 1e-5,
                        173, mode, dtype, fused=recurrence == 1)
    assert abs(res.iters - it_f) <= 1 and abs(res.adj_iters - it_a) <= 1
    if recurrence == 1:  # and against the reference iteration's own counts (include/pcg.cuh:118-208)
        _, it_fr, *_ = O.pcg(S.astype(dtype), P.astype(dtype), g.astype(dtype), np.zeros(n * N, dtype), n, N, 1e-5,
                             173, mode, dtype, fused=False)
        _, it_ar, *_ = O.pcg(S.astype(dtype), P.astype(dtype), glam.astype(dtype), np.zeros(n * N, dtype), n, N,
                             1e-5, 173, mode, dtype, fused=False)
        assert abs(res.iters - it_fr) <= 1 and abs(res.adj_iters - it_ar) <= 1, (res.iters, it_fr, res.adj_iters,
                                                                                  it_ar)
    # g_S is the masked, symmetrised outer product of the returned vectors
    G = -0.5 * (np.outer(gg, lam) + np.outer(lam, gg))
    expect = O.dense_to_cbtd(G, n, N)
    mask = O.pattern_mask(n, N)
    assert np.allclose(gS, expect * mask, rtol=1e-5 if dtype == np.float32 else 1e-12,
                       atol=1e-6 * np.abs(expect).max())
    assert not gS.reshape(N, 3, n * n)[0, 0].any() and not gS.reshape(N, 3, n * n)[-1, 2].any()
    if mode != 1:
        return
    # against the exact FP64 gradient (block-LU solves): both solves stop at a 1e-5
    # relative residual, so the gradient is good to that order
    x = O.block_solve(S, g, n, N)
    gg_x = O.block_solve(S, glam, n, N)
    gS_x = O.dense_to_cbtd(-0.5 * (np.outer(gg_x, x) + np.outer(x, gg_x)), n, N) * mask
    print(f"N={N} {dtype.__name__} rec={recurrence}: iters {res.iters}/{res.adj_iters} "
          f"g_gamma err vs exact {rel(gg, gg_x):.2e}, g_S err {rel(gS, gS_x):.2e}")
    assert rel(gg, gg_x) <= 1e-4 and rel(gS, gS_x) <= 1e-4


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("recurrence", [0, 1])
def test_adjoint_every_state_size(gpu, dtype, recurrence):
    """Forward + adjoint solves of every stateSize 1..32 at N in {2, 3, 5, 24}
    (fused and n <= 14 with N >= 4: both solves in one loop on the register
    kernel; else the LDS kernel, one after the other) against the oracle: both
    counts within one of the same-precision oracle's (both forms for the fused
    recurrence), lambda and g_gamma at the FP64 iterate of the same count, g_S by
    the formula.  FP32 to a relative residual of 3e-5 (above the FP32 floor of
    the smallest systems).  FP64 fused requests of n >= 25 at N in {3, 5}, and of
    n = 32 beyond N = 2, have no LDS layout with 2 knots per workgroup and run the
    reference iteration."""
    import gbdpcg
    tol = 3e-5 if dtype == np.float32 else 1e-9
    # FP64: as for single solves of every size, the summation order alone moves
    # the iterate of the least well-conditioned small systems by ~1e-10
    lam_tol = 1e-4 if dtype == np.float32 else 1e-9
    for n in range(1, 33):
        for N in (2, 3, 5, 24):
            S, P, g, glam = system(n, N, 6000 + 37 * n + N, dtype)
            lam, gg, gS, res = run_adjoint(S, P, g, glam, n, N, dtype, tol=tol, max_iter=400, exit_mode=1,
                                           recurrence=recurrence)
            assert res.status == 0 and res.adj_status == 0, (n, N, res.status, res.adj_status)
            if recurrence == 0:
                ran = False
            elif n in (2, 6, 12, 14, 28) and N >= 4:
                ran = True
            else:
                G = N // 2
                ran = gbdpcg.lds_bytes(dtype, n, -(-N // G), G) <= 160 * 1024
            for rhs, its, x in ((g, res.iters, lam), (glam, res.adj_iters, gg)):
                z = np.zeros(n * N, dtype)
                forms = (True, False) if recurrence == 1 else (False,)
                for f in forms:
                    _, it, *_ = O.pcg(S.astype(dtype), P.astype(dtype), rhs.astype(dtype), z, n, N, tol, 400, 1,
                                      dtype, fused=f)
                    assert abs(its - it) <= 1, (n, N, its, it, f)
                ref, it_ref, *_ = O.pcg(S, P, rhs, np.zeros(n * N), n, N, 0.0, its, 1, np.float64, fused=ran)
                assert it_ref == its
                assert rel(x, ref) <= lam_tol, (n, N, rel(x, ref))
            expect = O.dense_to_cbtd(-0.5 * (np.outer(gg, lam) + np.outer(lam, gg)), n, N) * O.pattern_mask(n, N)
            assert np.allclose(gS, expect, rtol=1e-5 if dtype == np.float32 else 1e-12,
                               atol=1e-6 * max(np.abs(expect).max(), 1e-300)), (n, N)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kernel", [2, 3])
def test_paired_every_horizon(gpu, dtype, kernel):
    """Forward + adjoint in one loop on the register layouts (column chunks: n in
    {2, 6, 12, 14}; row broadcast: n in {12, 14}) at every horizon N from 4 to
    40, against the oracle: both counts within one of both forms' same-precision
    counts, lambda and g_gamma at the FP64 iterates of those counts, g_S by the
    formula."""
    tol = 3e-5 if dtype == np.float32 else 1e-9
    bar = 1e-4 if dtype == np.float32 else 1e-9
    sizes = (2, 6, 12, 14) if kernel == 2 else (12, 14)
    for n in sizes:
        for N in range(4, 41):
            S, P, g, glam = system(n, N, 7500 + 41 * n + N, dtype)
            lam, gg, gS, res = run_adjoint(S, P, g, glam, n, N, dtype, tol=tol, max_iter=400, exit_mode=1,
                                           recurrence=1, kernel=kernel)
            assert res.status == 0 and res.adj_status == 0, (n, N, res.status, res.adj_status)
            for rhs, its, x in ((g, res.iters, lam), (glam, res.adj_iters, gg)):
                for f in (True, False):
                    _, it, *_ = O.pcg(S.astype(dtype), P.astype(dtype), rhs.astype(dtype), np.zeros(n * N, dtype), n,
                                      N, tol, 400, 1, dtype, fused=f)
                    assert abs(its - it) <= 1, (n, N, its, it, f)
                ref, *_ = O.pcg(S, P, rhs, np.zeros(n * N), n, N, 0.0, its, 1, np.float64, fused=True)
                assert rel(x, ref) <= bar, (n, N, rel(x, ref))
            expect = O.dense_to_cbtd(-0.5 * (np.outer(gg, lam) + np.outer(lam, gg)), n, N) * O.pattern_mask(n, N)
            assert np.allclose(gS, expect, rtol=1e-5 if dtype == np.float32 else 1e-12,
                               atol=1e-6 * np.abs(expect).max()), (n, N)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kernel", [1, 2, 3])
def test_sides_stop_independently(gpu, dtype, kernel):
    """Forward + adjoint, fused recurrence (the register kernels run both solves in
    one loop): each side stops on its own test, keeps its own status and count,
    and does not disturb the other."""
    import gbdpcg
    n, N = 14, 64
    S, P, g, glam = system(n, N, 4321, dtype)
    tol = 1e-5 if dtype == np.float32 else 1e-10
    bar = 1e-4 if dtype == np.float32 else 1e-10
    kw = dict(tol=tol, max_iter=173, exit_mode=1, recurrence=1, kernel=kernel)
    zero = np.zeros(n * N)

    def oracle_iters(rhs):
        return O.pcg(S.astype(dtype), P.astype(dtype), rhs.astype(dtype), np.zeros(n * N, dtype), n, N, tol, 173, 1,
                     dtype, fused=True)[1]

    # g_lambda = 0: the adjoint is converged at iteration 0 while the forward solve runs
    lam_a, gg, gS, res = run_adjoint(S, P, g, zero, n, N, dtype, **kw)
    assert (res.adj_iters, res.adj_status) == (0, gbdpcg.CONVERGED) and not gg.any() and not gS.any()
    assert res.status == gbdpcg.CONVERGED and abs(res.iters - oracle_iters(g)) <= 1
    assert rel(lam_a, O.pcg(S, P, g, zero, n, N, 0.0, res.iters)[0]) <= bar
    # gamma = 0: the forward solve is converged at iteration 0 while the adjoint runs
    lam, gg, gS, res = run_adjoint(S, P, zero, glam, n, N, dtype, **kw)
    assert (res.iters, res.status) == (0, gbdpcg.CONVERGED) and not lam.any() and not gS.any()
    assert res.adj_status == gbdpcg.CONVERGED and abs(res.adj_iters - oracle_iters(glam)) <= 1
    assert rel(gg, O.pcg(S, P, glam, zero, n, N, 0.0, res.adj_iters)[0]) <= bar
    # a NaN in g_lambda: the adjoint ends NONFINITE, the forward solve is unchanged
    bad = glam.copy()
    bad[7] = np.nan
    lam_c, gg, gS, res = run_adjoint(S, P, g, bad, n, N, dtype, **kw)
    assert res.adj_status == gbdpcg.NONFINITE and res.status == gbdpcg.CONVERGED
    assert np.array_equal(lam_c, lam_a)
    # max_iter = 5: both sides stop there, each at the oracle's 5-iteration iterate
    lam, gg, gS, res = run_adjoint(S, P, g, glam, n, N, dtype, **dict(kw, max_iter=5))
    assert (res.iters, res.status, res.adj_iters, res.adj_status) == (5, gbdpcg.MAX_ITER, 5, gbdpcg.MAX_ITER)
    assert rel(lam, O.pcg(S, P, g, zero, n, N, 0.0, 5)[0]) <= bar
    assert rel(gg, O.pcg(S, P, glam, zero, n, N, 0.0, 5)[0]) <= bar


@pytest.mark.parametrize("recurrence", [0, 1])
def test_adjoint_only_entry(gpu, recurrence):
    """forward = 0: lambda is an input; same g_gamma/g_S as the fused launch."""
    n, N, dtype = 14, 32, np.float64
    S, P, g, glam = system(n, N, 77, dtype)
    lam, gg, gS, res = run_adjoint(S, P, g, glam, n, N, dtype, tol=1e-10, max_iter=300, exit_mode=1,
                                   recurrence=recurrence)
    lam2, gg2, gS2, res2 = run_adjoint(S, P, g, glam, n, N, dtype, forward=False, lam_in=lam, tol=1e-10,
                                       max_iter=300, exit_mode=1, recurrence=recurrence)
    assert np.array_equal(lam2, lam)  # lambda untouched when it is an input
    assert res2.adj_iters == res.adj_iters
    assert np.allclose(gg2, gg, rtol=1e-12, atol=0) and np.allclose(gS2, gS, rtol=1e-12, atol=0)


def test_adjoint_fd_and_dense_fp64(gpu):
    """32 central finite differences (eps 1e-6) and the dense FP64 gradient."""
    n, N, dtype = 14, 16, np.float64
    S, P, g, glam = system(n, N, 4242, dtype)
    kw = dict(tol=1e-13, max_iter=600, exit_mode=1, recurrence=1)
    lam, gg, gS, res = run_adjoint(S, P, g, glam, n, N, dtype, **kw)
    gg_d, gS_d = O.adjoint_dense(S, O.dense_solve(S, g, n, N), glam, n, N)
    assert rel(gg, gg_d) < 1e-8 and rel(gS, gS_d) < 1e-8
    # finite differences of the exact FP64 solution map (a truncated iterative
    # solve would put its own truncation error into the difference quotient)
    def L(S_, g_):
        return float(glam @ O.dense_solve(S_, g_, n, N))

    rng = np.random.default_rng(5)
    mask = O.pattern_mask(n, N)
    eps = 1e-6
    for k in range(32):
        R = rng.normal(size=3 * n * n * N) * mask
        D = O.cbtd_to_dense(R, n, N)
        Dir = O.dense_to_cbtd(0.5 * (D + D.T), n, N)
        fd = (L(S + eps * Dir, g) - L(S - eps * Dir, g)) / (2 * eps)
        an = float(gS @ Dir)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (k, fd, an)
    dgam = rng.normal(size=n * N)
    fd = (L(S, g + eps * dgam) - L(S, g - eps * dgam)) / (2 * eps)
    assert abs(fd - float(gg @ dgam)) <= 1e-6 * max(1.0, abs(fd))


def test_autograd_function(gpu):
    import torch
    from gbdpcg.autograd import pcg_solve
    n, N = 6, 12
    S, P, g, glam = system(n, N, 31, np.float64)
    tS = _t(S, np.float64).requires_grad_(True)
    tg = _t(g, np.float64).requires_grad_(True)
    tP = _t(P, np.float64)
    lam = pcg_solve(tS, tg, tP, n, N, tol=1e-13, max_iter=500)
    loss = (lam * _t(glam, np.float64)).sum()
    loss.backward()
    gg_d, gS_d = O.adjoint_dense(S, O.dense_solve(S, g, n, N), glam, n, N)
    assert rel(tg.grad.cpu().numpy(), gg_d) < 1e-9
    assert rel(tS.grad.cpu().numpy(), gS_d) < 1e-9
    assert torch.isfinite(tS.grad).all()


def test_adjoint_fd_n14_N128_fp64(gpu):
    """The forward + adjoint configuration (n=14, N=128) through its FP64 check
    path: the device gradient against the exact FP64 gradient (block-LU solves)
    and against 32 central finite differences (eps 1e-6) of the exact FP64
    solution map over random symmetric block-tridiagonal directions, plus one
    direction in gamma."""
    n, N, dtype = 14, 128, np.float64
    S, P, g, glam = system(n, N, 4128, dtype)
    lam, gg, gS, res = run_adjoint(S, P, g, glam, n, N, dtype, tol=1e-13, max_iter=800, exit_mode=1, recurrence=1)
    assert res.status == 0 and res.adj_status == 0
    mask = O.pattern_mask(n, N)
    x = O.block_solve(S, g, n, N)
    gg_x = O.block_solve(S, glam, n, N)
    gS_x = O.dense_to_cbtd(-0.5 * (np.outer(gg_x, x) + np.outer(x, gg_x)), n, N) * mask
    assert rel(gg, gg_x) < 1e-8 and rel(gS, gS_x) < 1e-8, (rel(gg, gg_x), rel(gS, gS_x))

    def L(S_, g_):
        return float(glam @ O.block_solve(S_, g_, n, N))

    rng = np.random.default_rng(128)
    eps = 1e-6
    worst = 0.0
    for k in range(32):
        R = rng.normal(size=3 * n * n * N) * mask
        # symmetric direction in cbtd storage: block (k, k+1) is the transpose of (k+1, k)
        Dir = O.dense_to_cbtd(0.5 * (O.cbtd_to_dense(R, n, N) + O.cbtd_to_dense(R, n, N).T), n, N)
        fd = (L(S + eps * Dir, g) - L(S - eps * Dir, g)) / (2 * eps)
        an = float(gS @ Dir)
        worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (k, fd, an)
    dgam = rng.normal(size=n * N)
    fd = (L(S, g + eps * dgam) - L(S, g - eps * dgam)) / (2 * eps)
    assert abs(fd - float(gg @ dgam)) <= 1e-6 * max(1.0, abs(fd))
    print(f"n=14 N=128 FP64: iters {res.iters}/{res.adj_iters}, worst FD rel mismatch {worst:.1e}")


def test_autograd_function_fp32(gpu):
    """The autograd wrapper on the FP32 path at n=14 (the forward + adjoint
    configuration's precision and size class) against the exact FP64 gradient."""
    import torch
    from gbdpcg.autograd import pcg_solve
    n, N = 14, 128
    S, P, g, glam = system(n, N, 3131, np.float64)
    tS = _t(S, np.float32).requires_grad_(True)
    tg = _t(g, np.float32).requires_grad_(True)
    lam = pcg_solve(tS, tg, _t(P, np.float32), n, N, tol=1e-5, max_iter=173)
    (lam * _t(glam, np.float32)).sum().backward()
    x = O.block_solve(S.astype(np.float32).astype(np.float64), g.astype(np.float32).astype(np.float64), n, N)
    gg_x = O.block_solve(S.astype(np.float32).astype(np.float64), glam.astype(np.float32).astype(np.float64), n, N)
    gS_x = O.dense_to_cbtd(-0.5 * (np.outer(gg_x, x) + np.outer(x, gg_x)), n, N) * O.pattern_mask(n, N)
    e_lam = rel(lam.detach().cpu().numpy().astype(np.float64), x)
    e_gg = rel(tg.grad.cpu().numpy().astype(np.float64), gg_x)
    e_gS = rel(tS.grad.cpu().numpy().astype(np.float64), gS_x)
    print(f"autograd FP32 n=14 N=128: lambda {e_lam:.1e}, g_gamma {e_gg:.1e}, g_S {e_gS:.1e}")
    assert e_lam <= 1e-4 and e_gg <= 1e-4 and e_gS <= 1e-4
    assert tS.grad.dtype == torch.float32 and torch.isfinite(tS.grad).all()


def test_autograd_raises_when_a_solve_does_not_converge(gpu):
    """An unconverged forward (max_iter too small) or a non-finite adjoint
    (NaN in the upstream gradient) raises instead of returning a wrong value."""
    from gbdpcg.autograd import pcg_solve
    n, N = 6, 12
    S, P, g, glam = system(n, N, 37, np.float64)
    tS, tg, tP = _t(S, np.float64).requires_grad_(True), _t(g, np.float64).requires_grad_(True), _t(P, np.float64)
    with pytest.raises(RuntimeError, match="MAX_ITER"):
        pcg_solve(tS, tg, tP, n, N, tol=1e-13, max_iter=2)
    lam = pcg_solve(tS, tg, tP, n, N, tol=1e-12, max_iter=500)
    bad = glam.copy()
    bad[3] = np.nan
    with pytest.raises(RuntimeError, match="NONFINITE"):
        (lam * _t(bad, np.float64)).sum().backward()
