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
