"""GPU tests at the shapes of the two large batched workloads:

* 512 independent LQR-Schur systems, n drawn from {6, 12, 14, 28}, m = n/2,
  N uniform in [16, 256], symmetric-stair Pinv, tol 1e-5, one autotuned plan;
* n=28, m=14, N=512 systems, forward + adjoint, FP32 and FP64, ten warm-started
  re-solves (lambda shifted one knot, gamma += 1e-2 N(0,1)) captured in one graph.

Full-size checks use size-independent properties (every system converged, the
recomputed FP64 residual of every returned lambda meets the tolerance, graph
replays equal the eager sequence bit for bit); a sample of systems is compared
with the oracle run for the same iteration count.
"""
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


def true_rel_resid(S_dev, g_dev, lam_dev, n, N):
    S = S_dev.cpu().numpy().astype(np.float64)
    g = g_dev.cpu().numpy().astype(np.float64)
    lam = lam_dev.cpu().numpy().astype(np.float64)
    return float(np.linalg.norm(g - O.bdmv(S, lam, n, N)) / np.linalg.norm(g))


def make_systems(specs, dtype, seed0, adjoint=False):
    import torch
    import gbdpcg
    systems = []
    for i, (n, N) in enumerate(specs):
        S, g = gbdpcg.lqr_generate(n, n // 2, N, seed0 + i, dtype=dtype)
        dS = _t(S, dtype)
        dP = torch.empty_like(dS)
        gbdpcg.precond_device(dS, dP, n, N, gbdpcg.PRECOND_SS)
        d = dict(S=dS, Pinv=dP, gamma=_t(g, dtype), lam=torch.zeros(n * N, dtype=dS.dtype, device="cuda"), n=n, N=N,
                 tol=1e-5, max_iter=173)
        if adjoint:
            d.update(glam=_t(gbdpcg.randn(n * N, seed0 + 5000 + i, dtype=np.float64), dtype),
                     ggamma=torch.zeros(n * N, dtype=dS.dtype, device="cuda"), gS=torch.zeros_like(dS),
                     adjoint=True)
        systems.append(d)
    return systems


def test_uneven_512(gpu):
    import gbdpcg
    rng = np.random.default_rng(3000)
    specs = [(int(rng.choice([6, 12, 14, 28])), int(rng.integers(16, 257))) for _ in range(512)]
    print(f"512 uneven systems: sum nN = {sum(n * N for n, N in specs)}", flush=True)
    systems = make_systems(specs, np.float32, 3000)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED, autotune=True)
    print("tuned:", b.tune_report, b.info(), flush=True)
    b.run()
    res = b.results()
    assert all(r.status == gbdpcg.CONVERGED for r in res)
    worst = max(true_rel_resid(s["S"], s["gamma"], s["lam"], s["n"], s["N"]) for s in systems)
    assert worst <= 2e-5, worst
    for i in range(0, 512, 37):  # sample against the oracle
        s, r = systems[i], res[i]
        n, N = s["n"], s["N"]
        S = s["S"].cpu().numpy()
        P = s["Pinv"].cpu().numpy()
        g = s["gamma"].cpu().numpy()
        _, it_same, *_ = O.pcg(S, P, g, np.zeros(n * N, np.float32), n, N, 1e-5, 173, O.EXIT_REL, np.float32,
                               fused=True)
        assert abs(r.iters - it_same) <= 1, (i, n, N, r.iters, it_same)
        lam_ref, *_ = O.pcg(S.astype(np.float64), P.astype(np.float64), g.astype(np.float64), np.zeros(n * N), n, N,
                            0.0, r.iters)
        assert rel(s["lam"].cpu().numpy(), lam_ref) <= 1e-4, (i, n, N)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_warm_started_sweep_graph(gpu, dtype):
    """Ten warm-started forward+adjoint re-solves of n=28, N=512 systems in one
    graph: replays equal the eager sequence; the last solve matches the oracle."""
    import torch
    import gbdpcg
    count, n, N, steps = 6, 28, 512, 10
    systems = make_systems([(n, N)] * count, dtype, 5100, adjoint=True)
    lam0 = [s["lam"].clone() for s in systems]
    gam0 = [s["gamma"].clone() for s in systems]

    def restore():
        for s, l, g in zip(systems, lam0, gam0):
            s["lam"].copy_(l)
            s["gamma"].copy_(g)

    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    # eager sequence; keep the inputs of the last solve for the oracle
    last_in = None
    for k in range(steps):
        b.warm_start(seed=900 + k, noise_scale=1e-2)
        if k == steps - 1:
            last_in = [(s["lam"].cpu().numpy().copy(), s["gamma"].cpu().numpy().copy()) for s in systems]
        b.run()
        res = b.results()
        assert all(r.status == 0 and r.adj_status == 0 for r in res), k
    eager = [(s["lam"].cpu().numpy().copy(), s["gamma"].cpu().numpy().copy(), s["ggamma"].cpu().numpy().copy(),
              s["gS"].cpu().numpy().copy()) for s in systems]
    eager_res = [(r.iters, r.adj_iters) for r in res]
    for s in systems:
        assert true_rel_resid(s["S"], s["gamma"], s["lam"], n, N) <= 2e-5
    # oracle on the last solve (forward and adjoint)
    for s, (l_in, g_in), r in list(zip(systems, last_in, res))[:2]:
        S = s["S"].cpu().numpy().astype(np.float64)
        P = s["Pinv"].cpu().numpy().astype(np.float64)
        lam_ref, *_ = O.pcg(S, P, g_in.astype(np.float64), l_in.astype(np.float64), n, N, 0.0, r.iters)
        tol = 1e-4 if dtype == np.float32 else 1e-10
        assert rel(s["lam"].cpu().numpy(), lam_ref) <= tol
        gg_ref, *_ = O.pcg(S, P, s["glam"].cpu().numpy().astype(np.float64), np.zeros(n * N), n, N, 0.0, r.adj_iters)
        assert rel(s["ggamma"].cpu().numpy(), gg_ref) <= tol

    # the same sequence captured in one graph
    restore()
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        for k in range(steps):
            b.warm_start(seed=900 + k, noise_scale=1e-2, stream=stream)
            b.run(stream=stream)
    for rep in range(2):
        restore()
        junk = torch.randn(1 << 16, device="cuda")
        junk.mul_(3.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert [(r.iters, r.adj_iters) for r in b.results()] == eager_res
        for s, e in zip(systems, eager):
            assert np.array_equal(s["lam"].cpu().numpy(), e[0])
            assert np.array_equal(s["gamma"].cpu().numpy(), e[1])
            assert np.array_equal(s["ggamma"].cpu().numpy(), e[2])
            assert np.array_equal(s["gS"].cpu().numpy(), e[3])


def _oracle_counts(S, P, rhs, lam0, n, N, dtype):
    """Same-precision oracle iteration counts of the fused and the reference
    form (include/pcg.cuh:118-208), relative residual 1e-5, max_iter 173."""
    dt = np.dtype(dtype)
    args = (S.astype(dt), P.astype(dt), rhs.astype(dt), lam0.astype(dt), n, N, 1e-5, 173, O.EXIT_REL, dt)
    return O.pcg(*args, fused=True)[1], O.pcg(*args, fused=False)[1]


def test_forward_adjoint_64_n14_N128(gpu):
    """The forward + adjoint batch at its own configuration: 64 systems n=14, m=7,
    N=128, FP32, g_lambda ~ N(0,1), one autotuned persistent launch.  Every system
    against the oracle: forward and adjoint counts within +-1 of the fused and of
    the reference-form oracle, lambda and g_gamma within 1e-4 of the FP64 oracle
    run for the same number of iterations, g_S equal to its formula on the
    returned vectors."""
    import gbdpcg
    n, N, B = 14, 128, 64
    systems = make_systems([(n, N)] * B, np.float32, 4000, adjoint=True)
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED, autotune=True)
    for d in systems:
        d["lam"].zero_()
    b.run()
    res = b.results()
    zero = np.zeros(n * N)
    mask = O.pattern_mask(n, N)
    worst = [0.0, 0.0, 0.0]
    for i, (d, r) in enumerate(zip(systems, res)):
        assert r.status == gbdpcg.CONVERGED and r.adj_status == gbdpcg.CONVERGED, i
        S = d["S"].cpu().numpy().astype(np.float64)
        P = d["Pinv"].cpu().numpy().astype(np.float64)
        g = d["gamma"].cpu().numpy().astype(np.float64)
        gl = d["glam"].cpu().numpy().astype(np.float64)
        fw = _oracle_counts(S, P, g, zero, n, N, np.float32)
        ad = _oracle_counts(S, P, gl, zero, n, N, np.float32)
        assert all(abs(r.iters - c) <= 1 for c in fw), (i, r.iters, fw)
        assert all(abs(r.adj_iters - c) <= 1 for c in ad), (i, r.adj_iters, ad)
        lam = d["lam"].cpu().numpy().astype(np.float64)
        gg = d["ggamma"].cpu().numpy().astype(np.float64)
        e_l = rel(lam, O.pcg(S, P, g, zero, n, N, 0.0, r.iters)[0])
        e_g = rel(gg, O.pcg(S, P, gl, zero, n, N, 0.0, r.adj_iters)[0])
        expect = O.dense_to_cbtd(-0.5 * (np.outer(gg, lam) + np.outer(lam, gg)), n, N) * mask
        gS = d["gS"].cpu().numpy().astype(np.float64)
        e_s = float(np.abs(gS - expect).max() / np.abs(expect).max())
        worst = [max(worst[0], e_l), max(worst[1], e_g), max(worst[2], e_s)]
        assert e_l <= 1e-4 and e_g <= 1e-4 and e_s <= 1e-5, (i, e_l, e_g, e_s)
    print(f"64 x n14 N128 fwd+adj: worst lambda {worst[0]:.1e}, g_gamma {worst[1]:.1e}, g_S {worst[2]:.1e}; "
          f"mean iters {np.mean([r.iters for r in res]):.1f}/{np.mean([r.adj_iters for r in res]):.1f}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mpc_sweep_256_n28_N512(gpu, dtype):
    """The large sweep at its own configuration: 256 systems n=28, m=14, N=512,
    forward + adjoint, ten warm-started re-solves (lambda shifted one knot,
    gamma += 1e-2 N(0,1)).  Every system: both solves converged and the
    recomputed FP64 residuals of lambda and g_gamma at most 2e-5 relative; a
    sample against the oracle on the last re-solve's own inputs; the ten
    re-solves captured in one graph replay the eager sequence bit for bit."""
    import torch
    import gbdpcg
    count, n, N, steps = 256, 28, 512, 10
    systems = make_systems([(n, N)] * count, dtype, 5100, adjoint=True)
    lam0 = [s["lam"].clone() for s in systems]
    gam0 = [s["gamma"].clone() for s in systems]
    b = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    print(f"sweep {np.dtype(dtype).name}: plan {b.info()}", flush=True)
    last_in = None
    for k in range(steps):
        b.warm_start(seed=900 + k, noise_scale=1e-2)
        if k == steps - 1:
            last_in = [s["lam"].cpu().numpy().copy() for s in systems]
        b.run()
    res = b.results()
    assert all(r.status == 0 and r.adj_status == 0 for r in res)
    worst_l = worst_g = 0.0
    for s in systems:
        S = s["S"].cpu().numpy().astype(np.float64)
        for x, rhs, w in ((s["lam"], s["gamma"], 0), (s["ggamma"], s["glam"], 1)):
            rhs64 = rhs.cpu().numpy().astype(np.float64)
            rr = float(np.linalg.norm(rhs64 - O.bdmv(S, x.cpu().numpy().astype(np.float64), n, N)) /
                       np.linalg.norm(rhs64))
            if w == 0:
                worst_l = max(worst_l, rr)
            else:
                worst_g = max(worst_g, rr)
    print(f"worst recomputed residual: lambda {worst_l:.2e}, g_gamma {worst_g:.2e}", flush=True)
    assert worst_l <= 2e-5 and worst_g <= 2e-5
    tol = 1e-4 if dtype == np.float32 else 1e-10
    zero = np.zeros(n * N)
    for i in (0, 85, 170, 255):
        s, r = systems[i], res[i]
        S = s["S"].cpu().numpy().astype(np.float64)
        P = s["Pinv"].cpu().numpy().astype(np.float64)
        g = s["gamma"].cpu().numpy().astype(np.float64)
        gl = s["glam"].cpu().numpy().astype(np.float64)
        l_in = last_in[i].astype(np.float64)
        fw = _oracle_counts(S, P, g, l_in, n, N, dtype)
        ad = _oracle_counts(S, P, gl, zero, n, N, dtype)
        assert all(abs(r.iters - c) <= 1 for c in fw), (i, r.iters, fw)
        assert all(abs(r.adj_iters - c) <= 1 for c in ad), (i, r.adj_iters, ad)
        assert rel(s["lam"].cpu().numpy(), O.pcg(S, P, g, l_in, n, N, 0.0, r.iters)[0]) <= tol
        assert rel(s["ggamma"].cpu().numpy(), O.pcg(S, P, gl, zero, n, N, 0.0, r.adj_iters)[0]) <= tol
    eager = [(s["lam"].cpu().numpy().copy(), s["ggamma"].cpu().numpy().copy()) for s in systems]
    eager_it = [(r.iters, r.adj_iters) for r in res]
    # the same ten re-solves in one graph
    for s, l, g in zip(systems, lam0, gam0):
        s["lam"].copy_(l)
        s["gamma"].copy_(g)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        for k in range(steps):
            b.warm_start(seed=900 + k, noise_scale=1e-2, stream=stream)
            b.run(stream=stream)
    graph.replay()
    torch.cuda.synchronize()
    assert [(r.iters, r.adj_iters) for r in b.results()] == eager_it
    for s, (l, gg) in zip(systems, eager):
        assert np.array_equal(s["lam"].cpu().numpy(), l) and np.array_equal(s["ggamma"].cpu().numpy(), gg)
    b.close()
    del systems
    torch.cuda.empty_cache()
