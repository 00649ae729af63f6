"""GPU tests of the autotuner (single solves and batch plans), the knots-per-
workgroup knob, and the batched warm start with its counter-based noise."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _t(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).astype(dtype))).to(
        "cuda", torch.float64 if dtype == np.float64 else torch.float32)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("recurrence", [0, 1])
@pytest.mark.parametrize("N", [64, 128])
def test_tune_single_then_solve(gpu, recurrence, N):
    """The tuner times its candidates (the untuned default among them), keeps the
    fastest, and later automatic solves of that shape use it and stay at parity."""
    import gbdpcg
    n, dtype = 14, np.float32
    S, g = gbdpcg.lqr_generate(n, 7, N, 700 + N, dtype=np.float64)
    P = O.precond(S, n, N, 2)
    kw = dict(tol=1e-5, max_iter=173, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=recurrence)
    dS, dP, dg = _t(S, dtype), _t(P, dtype), _t(g, dtype)
    rep = gbdpcg.tune_single(dS, dg, n, N, dP, **kw)
    assert rep["candidates"] >= 4 and rep["from_cache"] == 0
    assert 0 < rep["best_us"] <= rep["default_us"]
    assert rep["kernel"] in (gbdpcg.KERNEL_LDS, gbdpcg.KERNEL_REGISTER, gbdpcg.KERNEL_ROWCAST)
    lam = _t(np.zeros(n * N), dtype)
    r = gbdpcg.solve(dS, dg, lam, n, N, dP, **kw)
    assert r.status == gbdpcg.CONVERGED
    _, it_same, *_ = O.pcg(S.astype(dtype), P.astype(dtype), g.astype(dtype), np.zeros(n * N, dtype), n, N, 1e-5,
                           173, O.EXIT_REL, dtype, fused=recurrence == 1)
    assert abs(r.iters - it_same) <= 1
    lam_ref, *_ = O.pcg(S, P, g, np.zeros(n * N), n, N, 0.0, r.iters)
    assert rel(lam.cpu().numpy(), lam_ref) <= 1e-4
    # a second call returns the cached choice without timing (Pinv may be absent)
    rep2 = gbdpcg.tune_single(dS, dg, n, N, None, **kw)
    assert rep2["from_cache"] == 1 and rep2["candidates"] == 0
    assert (rep2["threads"], rep2["wgs"], rep2["kernel"]) == (rep["threads"], rep["wgs"], rep["kernel"])


def test_tune_single_fused_request_without_fused_layout(gpu):
    """FP64 n = 25, N = 5 has no LDS layout with 2 knots per workgroup: tuning a
    fused request tunes the reference iteration's layouts, which the solve of the
    same config then runs (at parity with the reference form), and a second call
    finds that choice in the cache."""
    import gbdpcg
    n, N, dtype = 25, 5, np.float64
    S, g = gbdpcg.lqr_generate(n, 12, N, 731, dtype=np.float64)
    P = O.precond(S, n, N, 2)
    kw = dict(tol=1e-9, max_iter=400, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    dS, dP, dg = _t(S, dtype), _t(P, dtype), _t(g, dtype)
    rep = gbdpcg.tune_single(dS, dg, n, N, dP, **kw)
    assert rep["from_cache"] == 0 and rep["candidates"] >= 1 and rep["kernel"] == gbdpcg.KERNEL_LDS
    lam = _t(np.zeros(n * N), dtype)
    r = gbdpcg.solve(dS, dg, lam, n, N, dP, **kw)
    assert r.status == gbdpcg.CONVERGED
    lam_ref, it_ref, *_ = O.pcg(S, P, g, np.zeros(n * N), n, N, 1e-9, 400, O.EXIT_REL, np.float64, fused=False)
    assert abs(r.iters - it_ref) <= 1
    assert rel(lam.cpu().numpy(), O.dense_solve(S, g, n, N)) <= 1e-7
    rep2 = gbdpcg.tune_single(dS, dg, n, N, dP, **kw)
    assert rep2["from_cache"] == 1 and rep2["wgs"] == rep["wgs"]


def batch_systems(specs, dtype, seed0=12000):
    import torch
    import gbdpcg
    systems, host = [], []
    for i, (n, N) in enumerate(specs):
        S, g = gbdpcg.lqr_generate(n, n // 2, N, seed0 + i, dtype=np.float64)
        dS = _t(S, dtype)
        dP = torch.empty_like(dS)
        gbdpcg.precond_device(dS, dP, n, N, gbdpcg.PRECOND_SS)
        systems.append(dict(S=dS, Pinv=dP, gamma=_t(g, dtype), lam=_t(np.zeros(n * N), dtype), n=n, N=N,
                            tol=1e-5, max_iter=173))
        host.append((S, g))
    return systems, host


def test_batch_autotune_and_cache(gpu):
    import gbdpcg
    rng = np.random.default_rng(8)
    specs = [(int(rng.choice([6, 12, 14, 28])), int(rng.integers(16, 129))) for _ in range(24)]
    systems, host = batch_systems(specs, np.float32)
    lam0 = [s["lam"].clone() for s in systems]
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED, autotune=True)
    rep = b.tune_report
    assert rep["candidates"] >= 2 and rep["from_cache"] == 0 and rep["best_us"] <= rep["default_us"]
    # lambda restored after tuning
    for s, l0 in zip(systems, lam0):
        assert bool((s["lam"] == l0).all())
    b.run()
    res = b.results()
    for s, (S, g), r in zip(systems, host, res):
        n, N = s["n"], s["N"]
        assert r.status == gbdpcg.CONVERGED
        P = s["Pinv"].cpu().numpy().astype(np.float64)
        lam_ref, *_ = O.pcg(S, P, g, np.zeros(n * N), n, N, 0.0, r.iters)
        assert rel(s["lam"].cpu().numpy(), lam_ref) <= 1e-4
    # same systems again: the cached choice, nothing timed
    b2 = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED, autotune=True)
    assert b2.tune_report["from_cache"] == 1 and b2.tune_report["candidates"] == 0
    assert b2.wgs() == b.wgs()
    # the cached choice includes the launch split (one launch, or one per kernel class)
    assert b2.info()["launches"] == b.info()["launches"]


@pytest.mark.parametrize("cap", [2, 4, 9])
def test_knots_per_wg_cap(gpu, cap):
    import gbdpcg
    specs = [(14, 64), (6, 37), (28, 50)]
    systems, host = batch_systems(specs, np.float32)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED, knots_per_wg=cap)
    w = b.wgs()
    for (n, N), g in zip(specs, w):
        assert -(-N // g) <= cap or g == N // 2  # a cap of 2 on odd N: two knots minimum
    b.run()
    assert all(r.status == gbdpcg.CONVERGED for r in b.results())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_warm_start(gpu, dtype):
    """lambda shifted by one knot (last kept), gamma += scale * counter normals of
    (seed, system index); adjoint-only systems untouched."""
    import gbdpcg
    specs = [(14, 32), (6, 17), (28, 5)]
    systems, host = batch_systems(specs, dtype)
    rng = np.random.default_rng(1)
    for s in systems:
        s["lam"].copy_(_t(rng.standard_normal(s["n"] * s["N"]), dtype))
    extra = dict(systems[0])
    extra.update(lam=_t(rng.standard_normal(14 * 32), dtype), gamma=systems[0]["gamma"].clone(),
                 glam=_t(rng.standard_normal(14 * 32), dtype), ggamma=_t(np.zeros(14 * 32), dtype),
                 gS=_t(np.zeros(3 * 14 * 14 * 32), dtype), forward=False, adjoint=True)
    systems.append(extra)
    before = [(s["lam"].cpu().numpy().copy(), s["gamma"].cpu().numpy().copy()) for s in systems]
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    b.warm_start(seed=77, noise_scale=1e-2)
    import torch
    torch.cuda.synchronize()
    for i, (s, (l0, g0)) in enumerate(zip(systems, before)):
        l1, g1 = s["lam"].cpu().numpy(), s["gamma"].cpu().numpy()
        if i == len(systems) - 1:
            assert np.array_equal(l1, l0) and np.array_equal(g1, g0)
            continue
        n = s["n"]
        expect = l0.copy()
        expect[:-n] = l0[n:]
        assert np.array_equal(l1, expect)
        z = gbdpcg.counter_randn(len(g0), 77, i, 1e-2, dtype=dtype)
        # host and device evaluate the same double expression (libm vs device math
        # may differ in the last bit of log/cos)
        assert np.allclose(g1, g0 + z, rtol=0, atol=4 * np.finfo(dtype).eps * (1 + np.abs(g0).max()))
    # counter normals look standard normal
    z = gbdpcg.counter_randn(200000, 5, 3, 1.0, dtype=np.float64)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
