"""GPU batched path: many independent systems of uneven (n, N) in one persistent
launch, per-system tolerances, forward + adjoint, repeated runs and hipGraph
replays.  Each system is checked against the oracle like a single solve.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _t(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", torch.float64 if dtype == np.float64 else torch.float32)


def make_batch(specs, dtype, adjoint=False, seed0=9000):
    import gbdpcg
    systems, host = [], []
    for i, (n, N, tol) in enumerate(specs):
        S, g = gbdpcg.lqr_generate(n, max(n // 2, 1), N, seed0 + i, dtype=np.float64)
        P = O.precond(S, n, N, 2)
        glam = gbdpcg.randn(n * N, seed0 + 500 + i, dtype=np.float64)
        d = dict(S=_t(S.astype(dtype), dtype), Pinv=_t(P.astype(dtype), dtype), gamma=_t(g.astype(dtype), dtype),
                 lam=_t(np.zeros(n * N), dtype), n=n, N=N, tol=tol, max_iter=173)
        if adjoint:
            d.update(glam=_t(glam.astype(dtype), dtype), ggamma=_t(np.zeros(n * N), dtype),
                     gS=_t(np.zeros(3 * n * n * N), dtype), adjoint=True)
        systems.append(d)
        host.append((S, P, g, glam))
    return systems, host


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def check(systems, host, results, dtype, recurrence, adjoint=False, mode=O.EXIT_REL):
    """Every system against the oracle: the iteration count within one of the
    oracle's at the same precision (both forms of the recurrence for the fused
    one), lambda (and g_gamma, g_S for adjoint systems) at the oracle's iterate
    of the same count.  mode: the exit rule the plan ran with."""
    tol = 1e-4 if dtype == np.float32 else 1e-10

    def counts(rhs, tol_sys):
        z = np.zeros(rhs.size, dtype)
        _, it_same, *_ = O.pcg(S.astype(dtype), P.astype(dtype), rhs.astype(dtype), z, n, N, tol_sys, 173, mode,
                               dtype, fused=recurrence == 1)
        its = [it_same]
        if recurrence == 1:  # the fused form against the reference iteration's own count as well
            _, it_refform, *_ = O.pcg(S.astype(dtype), P.astype(dtype), rhs.astype(dtype), z, n, N, tol_sys, 173,
                                      mode, dtype, fused=False)
            its.append(it_refform)
        return its

    for d, (S, P, g, glam), r in zip(systems, host, results):
        n, N = d["n"], d["N"]
        assert r.status == 0, (n, N, r.status)
        for it in counts(g, d["tol"]):
            assert abs(r.iters - it) <= 1, (n, N, r.iters, it)
        lam_ref, *_ = O.pcg(S, P, g, np.zeros(n * N), n, N, 0.0, r.iters)
        assert rel(d["lam"].cpu().numpy(), lam_ref) <= tol, (n, N)
        if adjoint:
            assert r.adj_status == 0
            for it in counts(glam, d["tol"]):
                assert abs(r.adj_iters - it) <= 1, (n, N, "adjoint", r.adj_iters, it)
            gg_ref, *_ = O.pcg(S, P, glam, np.zeros(n * N), n, N, 0.0, r.adj_iters)
            gg = d["ggamma"].cpu().numpy().astype(np.float64)
            assert rel(gg, gg_ref) <= tol
            lam = d["lam"].cpu().numpy().astype(np.float64)
            expect = O.dense_to_cbtd(-0.5 * (np.outer(gg, lam) + np.outer(lam, gg)), n, N)
            assert np.allclose(d["gS"].cpu().numpy(), expect, rtol=1e-5, atol=1e-6 * np.abs(expect).max())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("recurrence", [0, 1])
def test_uneven_batch(gpu, dtype, recurrence):
    import gbdpcg
    rng = np.random.default_rng(3)
    specs = [(int(rng.choice([6, 12, 14, 28])), int(rng.integers(16, 97)), float(rng.choice([1e-5, 1e-7])))
             for _ in range(24)]
    systems, host = make_batch(specs, dtype)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=recurrence)
    info = b.info()
    assert info["systems"] == 24 and info["grid"] > 0
    b.run()
    check(systems, host, b.results(), dtype, recurrence)


def every_size_specs(dtype):
    """One system of every stateSize 1..32, N drawn from 5..40.  FP32 to a relative
    residual of 3e-5, above the FP32 floor of the smallest systems."""
    rng = np.random.default_rng(32)
    tol = 3e-5 if dtype == np.float32 else 1e-9
    return [(n, int(rng.integers(5, 41)), tol) for n in range(1, 33)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("split", [False, True])
def test_batch_every_state_size(gpu, dtype, split):
    """Forward + adjoint systems of every stateSize 1..32 in one plan (fused
    recurrence): one launch, or one per kernel class (register rows of n <= 14,
    rows of n = 28, the LDS kernel for the sizes the register kernels do not take).
    Every system against the oracle."""
    import gbdpcg
    specs = every_size_specs(dtype)
    systems, host = make_batch(specs, dtype, adjoint=True, seed0=9800)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED, split_launches=split)
    assert b.info()["systems"] == len(specs)
    b.run()
    check(systems, host, b.results(), dtype, 1, adjoint=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("recurrence", [0, 1])
@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("split", [False, True])
def test_batch_breakdown_does_not_stall(gpu, dtype, recurrence, adjoint, split):
    """Systems that break down exactly (p'Sp = 0, or r'z = 0 with r != 0) or hold
    a NaN, in a plan with ordinary ones: they end with BREAKDOWN / NONFINITE after
    0 iterations (both sides when paired with an adjoint of the same right-hand
    side), and every other system of the launch converges at parity with the
    oracle."""
    import gbdpcg
    tol = 3e-5 if dtype == np.float32 else 1e-9
    ok_specs = [(14, 32, tol), (7, 9, tol), (2, 6, tol), (28, 8, tol)]
    systems, host = make_batch(ok_specs, dtype, adjoint=adjoint, seed0=9750)
    n, N = 2, 4
    g = np.zeros(n * N)
    g[0] = g[1] = 1.0
    bad = []
    for sdiag, pdiag in (((1.0, -1.0), (1.0, 1.0)), ((2.0, 3.0), (1.0, -1.0))):
        S = np.zeros((N, 3, n, n))
        P = np.zeros((N, 3, n, n))
        for k in range(N):
            S[k, 1] = np.diag(sdiag)
            P[k, 1] = np.diag(pdiag)
        d = dict(S=_t(S.reshape(-1), dtype), Pinv=_t(P.reshape(-1), dtype), gamma=_t(g, dtype),
                 lam=_t(np.zeros(n * N), dtype), n=n, N=N, tol=tol, max_iter=173)
        if adjoint:
            d.update(glam=_t(g, dtype), ggamma=_t(np.zeros(n * N), dtype), gS=_t(np.zeros(3 * n * n * N), dtype),
                     adjoint=True)
        bad.append(d)
    # and a NaN in one right-hand side entry: NONFINITE
    S, gq = gbdpcg.lqr_generate(n, 1, N, 9799, dtype=np.float64)
    gq[3] = np.nan
    d = dict(S=_t(S, dtype), Pinv=_t(O.precond(S, n, N, 2), dtype), gamma=_t(gq, dtype),
             lam=_t(np.zeros(n * N), dtype), n=n, N=N, tol=tol, max_iter=173)
    if adjoint:
        d.update(glam=_t(gq, dtype), ggamma=_t(np.zeros(n * N), dtype), gS=_t(np.zeros(3 * n * n * N), dtype),
                 adjoint=True)
    bad.append(d)
    b = gbdpcg.Batch(bad[:1] + systems + bad[1:], exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=recurrence,
                     split_launches=split)
    b.run()
    res = b.results()
    k = len(systems)
    for r, d, want in ((res[0], bad[0], gbdpcg.BREAKDOWN), (res[k + 1], bad[1], gbdpcg.BREAKDOWN),
                       (res[k + 2], bad[2], gbdpcg.NONFINITE)):
        assert r.status == want and r.iters == 0 and not d["lam"].cpu().numpy().any(), (r.status, want)
        if adjoint:
            assert r.adj_status == want and r.adj_iters == 0
    check(systems, host, res[1:k + 1], dtype, recurrence, adjoint=adjoint)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_n28_paired_every_horizon(gpu, dtype):
    """One plan of n = 28 forward + adjoint systems, one per horizon N from 4 to
    40 (the paired instances at 512 threads: FP32 one lane per row, FP64 with S in
    LDS): both counts within one of both forms' same-precision counts, lambda and
    g_gamma at the fused FP64 iterates of those counts (1e-4 / 1e-9), g_S by the
    formula."""
    import gbdpcg
    tol = 3e-5 if dtype == np.float32 else 1e-9
    bar = 1e-4 if dtype == np.float32 else 1e-9
    specs = [(28, N, tol) for N in range(4, 41)]
    systems, host = make_batch(specs, dtype, adjoint=True, seed0=9850)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    if dtype == np.float64:
        assert b.info()["lds_bytes"] > 140 * 1024  # S held in LDS
    b.run()
    for d, (S, P, g, glam), r in zip(systems, host, b.results()):
        n, N = d["n"], d["N"]
        assert r.status == 0 and r.adj_status == 0, (N, r.status, r.adj_status)
        lam = d["lam"].cpu().numpy().astype(np.float64)
        gg = d["ggamma"].cpu().numpy().astype(np.float64)
        for rhs, its, x in ((g, r.iters, lam), (glam, r.adj_iters, gg)):
            for f in (True, False):
                _, it, *_ = O.pcg(S.astype(dtype), P.astype(dtype), rhs.astype(dtype), np.zeros(n * N, dtype), n, N,
                                  tol, 173, O.EXIT_REL, dtype, fused=f)
                assert abs(its - it) <= 1, (N, its, it, f)
            ref, *_ = O.pcg(S, P, rhs, np.zeros(n * N), n, N, 0.0, its, O.EXIT_REL, np.float64, fused=True)
            assert rel(x, ref) <= bar, (N, rel(x, ref))
        expect = O.dense_to_cbtd(-0.5 * (np.outer(gg, lam) + np.outer(lam, gg)), n, N)
        assert np.allclose(d["gS"].cpu().numpy(), expect, rtol=1e-5 if dtype == np.float32 else 1e-12,
                           atol=1e-6 * np.abs(expect).max()), N


def test_random_plans(gpu):
    """60 seeded random plans (1-40 systems each: n in {1, 2, 3, 6, 7, 12, 14, 20,
    28}, N in 1..96, forward-only / forward + adjoint / adjoint-only, some capped
    at 4 iterations; FP32 or FP64, either recurrence, one launch or split), every
    side of every system against the oracle: a binding cap ends in MAX_ITER after
    exactly the cap; otherwise CONVERGED with the count within one of each form's
    same-precision count, except where those two counts differ by more than one
    themselves (an FP32 fused solve at the precision floor), which is held to
    convergence only; the result within 1e-4 / 1e-9 of the FP64 iterate of either
    form at its count; g_S by the formula."""
    import gbdpcg
    rng = np.random.default_rng(5)
    for pi in range(60):
        dtype = [np.float32, np.float64][int(rng.integers(0, 2))]
        rec = int(rng.integers(0, 2))
        split = bool(rng.integers(0, 2))
        count = int(rng.integers(1, 41))
        tol = 3e-5 if dtype == np.float32 else 1e-9
        systems, meta = [], []
        for i in range(count):
            n = int(rng.choice([1, 2, 3, 6, 7, 12, 14, 20, 28]))
            N = int(rng.integers(1, 97))
            kind = ["fwd", "both", "adj"][int(rng.integers(0, 3))]
            mx = int(rng.choice([173, 173, 173, 4]))
            seed = int(rng.integers(0, 1 << 30))
            S, g = gbdpcg.lqr_generate(n, max(n // 2, 1), N, seed, dtype=np.float64)
            P = O.precond(S, n, N, 2)
            glam = gbdpcg.randn(n * N, seed + 7, dtype=np.float64)
            lam_in = O.dense_solve(S, g, n, N) if kind == "adj" else np.zeros(n * N)
            d = dict(S=_t(S.astype(dtype), dtype), Pinv=_t(P.astype(dtype), dtype), gamma=_t(g.astype(dtype), dtype),
                     lam=_t(lam_in.astype(dtype), dtype), n=n, N=N, tol=tol, max_iter=mx, forward=kind != "adj",
                     adjoint=kind != "fwd")
            if kind != "fwd":
                d.update(glam=_t(glam.astype(dtype), dtype), ggamma=_t(np.zeros(n * N), dtype),
                         gS=_t(np.zeros(3 * n * n * N), dtype))
            systems.append(d)
            meta.append((n, N, kind, mx, S, P, g, glam))
        b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=rec, split_launches=split)
        b.run()

        def side(case, S, P, rhs, n, N, mx, its, st, x):
            z = np.zeros(n * N, dtype)
            forms = (True, False) if rec else (False,)
            counts = [O.pcg(S.astype(dtype), P.astype(dtype), rhs.astype(dtype), z, n, N, tol, mx, O.EXIT_REL, dtype,
                            fused=f)[1:3] for f in forms]
            if counts[0][1] == O.MAX_ITER:
                assert st == gbdpcg.MAX_ITER and its == mx, case + (its, st, counts)
            else:
                assert st == gbdpcg.CONVERGED, case + (its, st, counts)
                floor = len(counts) == 2 and abs(counts[0][0] - counts[1][0]) > 1
                assert floor or all(abs(its - c) <= 1 for c, _ in counts), case + (its, counts)
            errs = [rel(x, O.pcg(S, P, rhs, np.zeros(n * N), n, N, 0.0, its, O.EXIT_REL, np.float64, fused=f)[0])
                    for f in (True, False)]
            assert min(errs) <= (1e-4 if dtype == np.float32 else 1e-9), case + tuple(errs)

        for i, (d, (n, N, kind, mx, S, P, g, glam), r) in enumerate(zip(systems, meta, b.results())):
            case = (pi, i, dtype.__name__, rec, split, n, N, kind, mx)
            lam = d["lam"].cpu().numpy().astype(np.float64)
            if kind != "adj":
                side(case, S, P, g, n, N, mx, r.iters, r.status, lam)
            if kind != "fwd":
                gg = d["ggamma"].cpu().numpy().astype(np.float64)
                side(case + ("adjoint",), S, P, glam, n, N, mx, r.adj_iters, r.adj_status, gg)
                expect = O.dense_to_cbtd(-0.5 * (np.outer(gg, lam) + np.outer(lam, gg)), n, N) * O.pattern_mask(n, N)
                assert np.allclose(d["gS"].cpu().numpy(), expect, rtol=1e-5 if dtype == np.float32 else 1e-12,
                                   atol=1e-6 * max(np.abs(expect).max(), 1e-300)), case
        b.close()


def test_batch_results_device(gpu):
    """gbdpcg_batch_results_device: the plan's device-resident result records
    (what a captured graph's later kernels can read) hold the same bytes as the
    host copy from gbdpcg_batch_results."""
    import ctypes
    import torch
    import gbdpcg
    specs = [(14, 32, 1e-5), (6, 20, 1e-5), (28, 16, 1e-5), (7, 9, 1e-5)]
    systems, host = make_batch(specs, np.float32, adjoint=True, seed0=9650)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    b.run()
    torch.cuda.synchronize()
    ptr = gbdpcg.lib().gbdpcg_batch_results_device(b._h)
    assert ptr
    nbytes = len(specs) * ctypes.sizeof(gbdpcg.Result)
    buf = (ctypes.c_uint8 * nbytes)()
    # the HIP runtime the library itself links (torch may bundle another copy,
    # which does not know this allocation)
    hip = ctypes.CDLL("libamdhip64.so.7")
    assert hip.hipMemcpy(buf, ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes), 2) == 0  # device to host
    res = b.results()
    assert bytes(buf) == b"".join(bytes(r) for r in res)
    check(systems, host, res, np.float32, 1, adjoint=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("split", [False, True])
def test_batch_per_system_wgs(gpu, dtype, split):
    """Per-system workgroup counts (the system's `wgs`, the reference's
    pcg_grid.x) are honoured in one plan (fused recurrence): one knot per
    workgroup runs the reference iteration, the others the fused one; every
    system against the oracle."""
    import gbdpcg
    tol = 3e-5 if dtype == np.float32 else 1e-9
    req = [(7, 12, 1), (7, 12, 3), (7, 12, 12), (14, 32, 8), (14, 32, 4), (14, 40, 0), (6, 20, 10), (28, 16, 8)]
    systems, host = make_batch([(n, N, tol) for n, N, _ in req], dtype, seed0=9600)
    for d, (_, _, w) in zip(systems, req):
        d["wgs"] = w
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED, split_launches=split)
    got = b.wgs()
    for (n, N, w), g in zip(req, got):
        assert g == w or (w == 0 and 1 <= g <= N), (n, N, w, g)
    b.run()
    check(systems, host, b.results(), dtype, 1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("split", [False, True])
def test_batch_kinds_and_limits(gpu, dtype, split):
    """One plan (fused recurrence) mixing forward-only, forward + adjoint and
    adjoint-only systems (the last with a given lambda), per-system max_iter caps
    far below convergence (those sides end in MAX_ITER after exactly max_iter
    iterations, at the oracle's iterate of that count) and horizons down to N = 1
    (one knot: the reference iteration).  FP64 lambda to 1e-9: the summation order
    alone moves the iterate of small, less well-conditioned systems by ~1e-10."""
    import gbdpcg
    tol = 3e-5 if dtype == np.float32 else 1e-9
    lam_tol = 1e-4 if dtype == np.float32 else 1e-9
    specs = [(14, 32, "both", 173), (14, 32, "both", 5), (14, 40, "fwd", 4), (14, 24, "adj", 173),
             (6, 3, "both", 173), (6, 2, "fwd", 173), (6, 1, "both", 173), (2, 1, "fwd", 173),
             (28, 16, "both", 173), (28, 9, "both", 6), (7, 12, "adj", 173), (7, 3, "fwd", 2),
             (12, 4, "both", 173), (2, 5, "adj", 3)]
    systems, host = [], []
    for i, (n, N, kind, mx) in enumerate(specs):
        S, g = gbdpcg.lqr_generate(n, max(n // 2, 1), N, 9900 + i, dtype=np.float64)
        P = O.precond(S, n, N, 2)
        glam = gbdpcg.randn(n * N, 9950 + i, dtype=np.float64)
        lam_in = O.dense_solve(S, g, n, N) if kind == "adj" else np.zeros(n * N)
        d = dict(S=_t(S.astype(dtype), dtype), Pinv=_t(P.astype(dtype), dtype), gamma=_t(g.astype(dtype), dtype),
                 lam=_t(lam_in.astype(dtype), dtype), n=n, N=N, tol=tol, max_iter=mx, forward=kind != "adj",
                 adjoint=kind != "fwd")
        if kind != "fwd":
            d.update(glam=_t(glam.astype(dtype), dtype), ggamma=_t(np.zeros(n * N), dtype),
                     gS=_t(np.zeros(3 * n * n * N), dtype))
        systems.append(d)
        host.append((S, P, g, glam, lam_in))
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED, split_launches=split)
    assert b.info()["systems"] == len(specs)
    b.run()

    def side(S, P, rhs, n, N, mx, its, status, out):
        z = np.zeros(n * N, dtype)
        for f in (True, False):
            _, it, st, *_ = O.pcg(S.astype(dtype), P.astype(dtype), rhs.astype(dtype), z, n, N, tol, mx, O.EXIT_REL,
                                  dtype, fused=f)
            if mx < 10:  # the cap binds far below convergence
                assert st == gbdpcg.MAX_ITER and status == gbdpcg.MAX_ITER and its == mx, (n, N, its, st, status)
            else:
                assert status == gbdpcg.CONVERGED and abs(its - it) <= 1, (n, N, its, it, f)
        ref, it_ref, *_ = O.pcg(S, P, rhs, np.zeros(n * N), n, N, 0.0, its, O.EXIT_REL, np.float64, fused=N >= 2)
        assert it_ref == its
        assert rel(out, ref) <= lam_tol, (n, N, rel(out, ref))

    for (n, N, kind, mx), d, (S, P, g, glam, lam_in), r in zip(specs, systems, host, b.results()):
        lam = d["lam"].cpu().numpy().astype(np.float64)
        if kind == "adj":
            assert np.array_equal(lam, lam_in.astype(dtype).astype(np.float64))  # lambda is an input here
        else:
            side(S, P, g, n, N, mx, r.iters, r.status, lam)
        if kind != "fwd":
            gg = d["ggamma"].cpu().numpy().astype(np.float64)
            side(S, P, glam, n, N, mx, r.adj_iters, r.adj_status, gg)
            expect = O.dense_to_cbtd(-0.5 * (np.outer(gg, lam) + np.outer(lam, gg)), n, N) * O.pattern_mask(n, N)
            assert np.allclose(d["gS"].cpu().numpy(), expect, rtol=1e-5 if dtype == np.float32 else 1e-12,
                               atol=1e-6 * np.abs(expect).max()), (n, N)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("split", [False, True])
def test_kernel_classes_in_one_plan(gpu, dtype, adjoint, split):
    """Rows of n <= 14, rows of n = 28 and a stateSize the register kernels do
    not take (n = 7) in one plan (fused recurrence): one persistent launch (the
    LDS kernel for all), or with split_launches one launch per kernel class,
    back to back.  Every system against the oracle, forward and adjoint; then
    the same plan without n = 7."""
    import gbdpcg
    tol = 1e-5 if dtype == np.float32 else 1e-9
    specs = [(6, 24, tol), (14, 40, tol), (28, 20, tol), (7, 18, tol), (14, 9, tol), (28, 33, tol), (7, 5, tol)]
    systems, host = make_batch(specs, dtype, adjoint=adjoint, seed0=9700)
    kw = dict(exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED, split_launches=split)
    b = gbdpcg.Batch(systems, **kw)
    info = b.info()
    assert info["launches"] == (3 if split else 1) and info["systems"] == len(specs)
    b.run()
    check(systems, host, b.results(), dtype, 1, adjoint=adjoint)
    # the register kernel cannot take n = 7: forcing it is declined
    with pytest.raises(gbdpcg.GbdpcgError) as e:
        gbdpcg.Batch(systems, kernel=gbdpcg.KERNEL_REGISTER, **kw)
    assert e.value.code == gbdpcg.ERR_UNSUPPORTED
    keep = [i for i, s in enumerate(specs) if s[0] != 7]
    sub, sub_host = [systems[i] for i in keep], [host[i] for i in keep]
    for d in sub:
        for key in ("lam", "ggamma", "gS"):
            if key in d:
                d[key].zero_()
    b2 = gbdpcg.Batch(sub, **kw)
    # one launch for everything, except FP64 mixing n = 28 with smaller rows
    # (no runtime-n instance holds FP64 rows of n = 28: the LDS kernel runs it)
    assert b2.info()["launches"] == (2 if split else 1)
    b2.run()
    check(sub, sub_host, b2.results(), dtype, 1, adjoint=adjoint)


@pytest.mark.parametrize("recurrence", [0, 1])
def test_forward_adjoint_batch(gpu, recurrence):
    import gbdpcg
    specs = [(14, 32, 1e-5)] * 6 + [(6, 20, 1e-6)] * 3
    systems, host = make_batch(specs, np.float32, adjoint=True)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=recurrence)
    b.run()
    check(systems, host, b.results(), np.float32, recurrence, adjoint=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [12, 14])
@pytest.mark.parametrize("kernel", [2, 3])
def test_register_layouts_forward_adjoint(gpu, dtype, n, kernel):
    """Batches of one stateSize, forward + adjoint, forced onto each register
    layout (column chunks; row broadcast), uneven N and workgroups of 1..30 knots."""
    import gbdpcg
    tol = 1e-5 if dtype == np.float32 else 1e-10
    specs = [(n, N, tol) for N in (4, 16, 33, 64, 97, 128, 200)]
    systems, host = make_batch(specs, dtype, adjoint=True, seed0=9300 + n)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=1, kernel=kernel)
    b.run()
    check(systems, host, b.results(), dtype, 1, adjoint=True)


@pytest.mark.parametrize("variant", ["paired", "mixed", "budget"])
def test_fp64_n28_forward_adjoint(gpu, variant):
    """FP64 rows of n = 28, forward + adjoint: a plan of paired systems only runs
    the 512-thread instance that keeps S in LDS (from one rank for N = 5 to
    ranks of 7 knots); with one forward-only system in the plan, or with an LDS
    budget below what that instance needs, it runs the 256-thread
    one-lane-per-row instance instead.  Every system against the oracle."""
    import gbdpcg
    mixed = variant == "mixed"
    specs = [(28, N, 1e-10) for N in (5, 14, 40, 97)]
    systems, host = make_batch(specs, np.float64, adjoint=True, seed0=9700)
    if mixed:
        systems[1]["adjoint"] = False
    budget = 64 * 1024 if variant == "budget" else 0
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=1, lds_budget=budget)
    # the S chunks take 141 KB of LDS per workgroup; the other instance ~26 KB
    lds = b.info()["lds_bytes"]
    assert (lds > 140 * 1024) if variant == "paired" else (lds < 64 * 1024), lds
    b.run()
    res = b.results()
    if mixed:
        check(systems[1:2], host[1:2], res[1:2], np.float64, 1)
        keep = [0, 2, 3]
        check([systems[i] for i in keep], [host[i] for i in keep], [res[i] for i in keep], np.float64, 1,
              adjoint=True)
    else:
        check(systems, host, res, np.float64, 1, adjoint=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gS_unaligned_output(gpu, dtype):
    """g_S is stored 16 bytes at a time where its address allows: a g_S buffer
    offset by one element from a 16-byte boundary takes the element-wise
    stores, with the same values (checked like every other system)."""
    import torch
    import gbdpcg
    tol = 1e-5 if dtype == np.float32 else 1e-10
    specs = [(28, 40, tol), (28, 33, tol), (14, 32, tol)]
    systems, host = make_batch(specs, dtype, adjoint=True, seed0=9900)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    for d in systems[:2]:
        n, N = d["n"], d["N"]
        d["gS"] = torch.zeros(3 * n * n * N + 1, dtype=tdt, device="cuda")[1:]
        assert d["gS"].data_ptr() % 16 != 0
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=1)
    b.run()
    check(systems, host, b.results(), dtype, 1, adjoint=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_reference_exit_rule(gpu, dtype):
    """Batch plans under the reference's exit rule |r'z| < tol
    (include/pcg.cuh:195): forward + adjoint systems of n in {6, 14, 28}, each
    side's count within one of the oracle's (both recurrence forms), lambda,
    g_gamma and g_S at the oracle's iterates."""
    import gbdpcg
    tol = 1e-5 if dtype == np.float32 else 1e-10
    specs = [(6, 24, tol), (14, 40, tol), (28, 33, tol), (14, 97, tol)]
    systems, host = make_batch(specs, dtype, adjoint=True, seed0=9800)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_ABS_ETA, recurrence=1)
    b.run()
    check(systems, host, b.results(), dtype, 1, adjoint=True, mode=O.EXIT_ABS)


def test_repeated_runs_and_small_lds_budget(gpu):
    """Runs continue the exchange epochs; a small LDS budget forces more workgroups
    per system and several rounds per workgroup."""
    import torch
    import gbdpcg
    specs = [(14, 64, 1e-5)] * 10 + [(28, 40, 1e-5)] * 4
    systems, host = make_batch(specs, np.float32)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=1, lds_budget=64 * 1024)
    info = b.info()
    assert info["lds_bytes"] <= 64 * 1024
    first = None
    for rep in range(3):
        for d in systems:
            d["lam"].zero_()
        b.run()
        res = b.results()
        lams = [d["lam"].cpu().numpy().copy() for d in systems]
        if first is None:
            first = (lams, [r.iters for r in res])
            check(systems, host, res, np.float32, 1)
        else:
            assert [r.iters for r in res] == first[1]
            for a, c in zip(lams, first[0]):
                assert np.array_equal(a, c)
    torch.cuda.synchronize()


def test_graph_capture_replay(gpu):
    """The batch launch captured in a graph and replayed, with eager work and
    synchronisations between replays; every replay matches the eager result."""
    import torch
    import gbdpcg
    specs = [(14, 48, 1e-5)] * 8
    systems, host = make_batch(specs, np.float32)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=1)
    b.run()
    eager = [d["lam"].cpu().numpy().copy() for d in systems]
    eager_iters = [r.iters for r in b.results()]
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    for d in systems:
        d["lam"].zero_()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        for d in systems:
            d["lam"].zero_()
        b.run(stream=s)
    for rep in range(4):
        junk = torch.randn(1 << 16, device="cuda")
        junk.mul_(2.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        res = b.results()
        assert [r.iters for r in res] == eager_iters
        for d, e in zip(systems, eager):
            assert np.array_equal(d["lam"].cpu().numpy(), e)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("adjoint", [False, True])
def test_padding_blocks_never_used(gpu, dtype, adjoint):
    """Batched plans (runtime-n and paired instances) with NaN in every system's
    padding blocks S_{0,-1}, S_{N-1,N} (and Pinv's) give bit-identical results
    to zero padding."""
    import gbdpcg
    tol = 1e-5 if dtype == np.float32 else 1e-10
    specs = [(14, 40, tol), (6, 17, tol), (12, 33, tol), (14, 8, tol)]
    out = []
    for fill in (0.0, np.nan):
        systems, _ = make_batch(specs, dtype, adjoint=adjoint, seed0=9400)
        for d in systems:
            n, N = d["n"], d["N"]
            bs = 3 * n * n
            for key in ("S", "Pinv"):
                d[key][:n * n] = fill
                d[key][(N - 1) * bs + 2 * n * n:N * bs] = fill
        plan = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
        plan.run()
        res = plan.results()
        assert all(r.status == gbdpcg.CONVERGED for r in res)
        out.append(([r.iters for r in res], [d["lam"].cpu().numpy() for d in systems],
                    [d["ggamma"].cpu().numpy() for d in systems] if adjoint else []))
        plan.close()
    assert out[0][0] == out[1][0]
    for a, b in zip(out[0][1] + out[0][2], out[1][1] + out[1][2]):
        assert np.array_equal(a, b)
