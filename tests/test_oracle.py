"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (a transposed
block, a wrong halo index, a sign, a dropped term in the recurrence, an
off-by-one iteration count) fails it, and none retypes the oracle's own formula:
  * layout: the reference example's S is symmetric only under the right cbtd
    reading; the block matvec equals a dense matmul of the independently
    expanded matrix; padding blocks are never read;
  * recurrence: FP64 iterates equal SciPy's preconditioned CG (an independent
    implementation) step by step; the recurrence residual equals the recomputed
    gamma - S lambda; exact invariance under scaling by a power of two;
  * iteration count: one iteration when the tolerance is huge, zero when
    max_iter is 0, one when Pinv = S^-1 (single block); the reference example
    converges in its dimension (6) in FP64;
  * solution: the block-tridiagonal direct solve and a converged PCG agree with
    numpy's dense solve; the reference example matches its closed form;
  * preconditioners: N = 1 gives S^-1; 1x1 blocks give the hand-derived
    symmetric-stair closed form; phi = 0 reduces SS to block Jacobi;
  * adjoint: central finite differences of lambda along symmetric
    block-tridiagonal perturbations and along gamma.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def readme():
    d = json.load(open(os.path.join(HERE, "golden", "readme_example.json")))
    S = np.array(d["S"], dtype=np.float64)
    P = np.where(S != 0, 1 / np.where(S != 0, S, 1), 0)
    return d, S, P, np.array(d["gamma"], dtype=np.float64)


def random_spd_cbtd(n, N, seed, off=0.3):
    """Symmetric positive definite block-tridiagonal matrix built densely, then
    compressed (independent of any solver code)."""
    rng = np.random.default_rng(seed)
    D = np.zeros((n * N, n * N))
    for k in range(N):
        B = rng.normal(size=(n, n))
        D[k * n:(k + 1) * n, k * n:(k + 1) * n] = B @ B.T + n * np.eye(n)
        if k > 0:
            C = off * rng.normal(size=(n, n))
            D[k * n:(k + 1) * n, (k - 1) * n:k * n] = C
            D[(k - 1) * n:k * n, k * n:(k + 1) * n] = C.T
    return O.dense_to_cbtd(D, n, N), D


# ---------------------------------------------------------------------------
# layout and block matvec
# ---------------------------------------------------------------------------

def test_readme_layout_is_symmetric():
    d, S, _, _ = readme()
    D = O.cbtd_to_dense(S, d["n"], d["N"])
    assert np.array_equal(D, D.T)
    # reading a block row as a row-major n x 3n matrix (a plausible layout
    # mistake) gives a non-symmetric matrix
    n, N = d["n"], d["N"]
    Db = np.zeros((n * N, n * N))
    for k in range(N):
        for r in range(n):
            for c in range(3 * n):
                j = (k - 1) * n + c
                if 0 <= j < n * N:
                    Db[k * n + r, j] = S[k * 3 * n * n + r * 3 * n + c]
    assert not np.array_equal(Db, Db.T)


@pytest.mark.parametrize("n,N", [(1, 1), (2, 2), (2, 3), (3, 5), (5, 4), (14, 6)])
def test_bdmv_matches_dense(n, N):
    rng = np.random.default_rng(n * 100 + N)
    A = rng.normal(size=3 * n * n * N)
    x = rng.normal(size=n * N)
    D = O.cbtd_to_dense(A, n, N)
    for dt, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
        y = O.bdmv(A, x, n, N, dt)
        assert np.allclose(y, D @ x, rtol=tol, atol=tol * np.abs(D).sum(1).max())


def test_bdmv_never_reads_padding():
    n, N = 3, 4
    rng = np.random.default_rng(1)
    A = rng.normal(size=3 * n * n * N)
    A4 = A.reshape(N, 3, n * n).copy()
    A4[0, 0] = np.nan
    A4[N - 1, 2] = np.nan
    y = O.bdmv(A4.reshape(-1), rng.normal(size=n * N), n, N)
    assert np.isfinite(y).all()


def test_cbtd_roundtrip():
    S, D = random_spd_cbtd(4, 5, 3)
    assert np.array_equal(O.dense_to_cbtd(O.cbtd_to_dense(S, 4, 5), 4, 5), S)


# ---------------------------------------------------------------------------
# PCG recurrence
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,N,seed", [(3, 6, 0), (6, 10, 1), (14, 8, 2)])
def test_iterates_match_scipy_pcg(n, N, seed):
    """Step-by-step agreement with SciPy's preconditioned CG (block-Jacobi M)."""
    S, D = random_spd_cbtd(n, N, seed)
    P = O.precond(S, n, N, 1)
    M = O.cbtd_to_dense(P, n, N)
    b = np.random.default_rng(seed + 7).normal(size=n * N)
    iterates = []
    spla.cg(D, b, x0=np.zeros(n * N), rtol=0.0, atol=0.0, maxiter=8, M=spla.aslinearoperator(M),
            callback=lambda xk: iterates.append(xk.copy()))
    assert len(iterates) >= 5
    for k, xk in enumerate(iterates, start=1):
        lam, it, st, *_ = O.pcg(S, P, b, np.zeros(n * N), n, N, 0.0, k)
        assert it == k
        assert np.allclose(lam, xk, rtol=1e-9, atol=1e-12 * np.abs(xk).max())


def test_fused_form_matches_reference_form():
    S, D = random_spd_cbtd(6, 12, 5)
    P = O.precond(S, 6, 12, 2)
    b = np.random.default_rng(3).normal(size=72)
    for k in (1, 3, 7, 12):
        a, ia, *_ = O.pcg(S, P, b, np.zeros(72), 6, 12, 0.0, k)
        f, if_, *_ = O.pcg(S, P, b, np.zeros(72), 6, 12, 0.0, k, fused=True)
        assert ia == if_ == k
        assert np.allclose(a, f, rtol=1e-9, atol=1e-12)


def test_recurrence_residual_is_true_residual():
    S, D = random_spd_cbtd(5, 9, 11)
    P = O.precond(S, 5, 9, 2)
    b = np.random.default_rng(4).normal(size=45)
    for k in (1, 4, 10):
        lam, it, st, eta, rr, _ = O.pcg(S, P, b, np.zeros(45), 5, 9, 0.0, k)
        r = b - D @ lam
        assert np.isclose(np.sqrt(rr), np.linalg.norm(r), rtol=1e-8, atol=1e-13)


def test_scaling_by_power_of_two_is_exact():
    S, D = random_spd_cbtd(4, 7, 21)
    P = O.precond(S, 4, 7, 2)
    b = np.random.default_rng(9).normal(size=28)
    lam1, it1, *_ = O.pcg(S, P, b, np.zeros(28), 4, 7, 1e-9, 100, O.EXIT_REL)
    lam2, it2, *_ = O.pcg(4 * S, P, 4 * b, np.zeros(28), 4, 7, 1e-9, 100, O.EXIT_REL)
    assert it1 == it2 and np.array_equal(lam1, lam2)


def test_iteration_count_semantics():
    S, D = random_spd_cbtd(3, 5, 2)
    P = O.precond(S, 3, 5, 2)
    b = np.random.default_rng(2).normal(size=15)
    lam0 = np.linspace(-1, 1, 15)
    # max_iter = 0: nothing happens, status MAX_ITER
    lam, it, st, *_ = O.pcg(S, P, b, lam0, 3, 5, 1e-9, 0)
    assert it == 0 and st == O.MAX_ITER and np.array_equal(lam, lam0)
    # the test comes after the update: a huge tolerance still costs one iteration
    lam, it, st, *_ = O.pcg(S, P, b, np.zeros(15), 3, 5, 1e300, 10)
    assert it == 1 and st == O.CONVERGED
    # zero right-hand side and guess: converged without iterating
    lam, it, st, *_ = O.pcg(S, P, np.zeros(15), np.zeros(15), 3, 5, 1e-9, 10)
    assert it == 0 and st == O.CONVERGED


def test_exact_preconditioner_single_block():
    S, D = random_spd_cbtd(6, 1, 8)
    P = O.dense_to_cbtd(np.linalg.inv(D), 6, 1)
    b = np.random.default_rng(8).normal(size=6)
    lam, it, st, *_ = O.pcg(S, P, b, np.zeros(6), 6, 1, 1e-12, 10, O.EXIT_REL)
    assert it == 1 and st == O.CONVERGED
    assert np.allclose(lam, np.linalg.solve(D, b), rtol=1e-12)


def test_readme_example_fp64():
    d, S, P, g = readme()
    lam, it, st, eta, rr, hist = O.pcg(S, P, g, np.zeros(6), 2, 3, d["tol"], d["max_iter"])
    exact = np.linalg.solve(O.cbtd_to_dense(S, 2, 3), g)
    assert st == O.CONVERGED and it == 6  # converges in the dimension of the system
    assert np.allclose(lam, exact, rtol=1e-10)
    assert np.allclose(exact, d["lambda_closed_form"], rtol=1e-8)


# ---------------------------------------------------------------------------
# direct solves
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,N", [(1, 1), (2, 3), (6, 9), (14, 12)])
def test_block_solve_matches_dense(n, N):
    S, D = random_spd_cbtd(n, N, n + N)
    b = np.random.default_rng(n).normal(size=n * N)
    assert np.allclose(O.block_solve(S, b, n, N), np.linalg.solve(D, b), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("N", [6, 16])
def test_block_solve_with_pivoting(N):
    """Blocks whose LU needs row interchanges at several steps (tiny leading
    diagonal, nonsymmetric): a solver that applies the interchanges out of order
    fails this, while diagonally dominant SPD blocks never pivot."""
    n = 5
    rng = np.random.default_rng(N)
    D = np.zeros((n * N, n * N))
    for k in range(N):
        B = rng.normal(size=(n, n))
        B[np.arange(n), np.arange(n)] *= 1e-3
        D[k * n:(k + 1) * n, k * n:(k + 1) * n] = B + 0.5 * np.roll(np.eye(n), 1, axis=1) * n
        if k > 0:
            D[k * n:(k + 1) * n, (k - 1) * n:k * n] = 0.2 * rng.normal(size=(n, n))
            D[(k - 1) * n:k * n, k * n:(k + 1) * n] = 0.2 * rng.normal(size=(n, n))
    S = O.dense_to_cbtd(D, n, N)
    b = rng.normal(size=n * N)
    assert np.allclose(O.block_solve(S, b, n, N), np.linalg.solve(D, b), rtol=1e-9, atol=1e-11)


def test_block_solve_lqr_system():
    import gbdpcg
    S, g = gbdpcg.lqr_generate(14, 7, 16, 2128, dtype=np.float64)
    D = O.cbtd_to_dense(S, 14, 16)
    x = O.block_solve(S, g, 14, 16)
    assert np.linalg.norm(D @ x - g) <= 1e-12 * np.linalg.norm(g)


def test_block_solve_negative_definite():
    d, S, _, g = readme()
    assert np.allclose(O.block_solve(S, g, 2, 3), d["lambda_closed_form"], rtol=1e-8)


def test_converged_pcg_matches_direct():
    S, D = random_spd_cbtd(14, 16, 4)
    P = O.precond(S, 14, 16, 2)
    b = np.random.default_rng(1).normal(size=14 * 16)
    lam, it, st, *_ = O.pcg(S, P, b, np.zeros(14 * 16), 14, 16, 1e-13, 500, O.EXIT_REL)
    assert st == O.CONVERGED
    assert np.allclose(lam, O.block_solve(S, b, 14, 16), rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------------------
# preconditioners (definition)
# ---------------------------------------------------------------------------

def test_precond_single_block_is_inverse():
    S, D = random_spd_cbtd(5, 1, 3)
    for kind in (1, 2):
        assert np.allclose(O.cbtd_to_dense(O.precond(S, 5, 1, kind), 5, 1), np.linalg.inv(D))


def test_symmetric_stair_scalar_closed_form():
    # S = [[a, b], [b, c]] with 1x1 blocks: SS = [[1/a, -b/(ac)], [-b/(ac), 1/c]]
    a, b, c = 3.0, -0.7, 2.5
    S = O.dense_to_cbtd(np.array([[a, b], [b, c]]), 1, 2)
    P = O.cbtd_to_dense(O.precond(S, 1, 2, 2), 1, 2)
    assert np.allclose(P, [[1 / a, -b / (a * c)], [-b / (a * c), 1 / c]])


def test_symmetric_stair_reduces_to_jacobi_without_coupling():
    S, D = random_spd_cbtd(4, 6, 5, off=0.0)
    assert np.allclose(O.precond(S, 4, 6, 2), O.precond(S, 4, 6, 1))


# ---------------------------------------------------------------------------
# adjoint
# ---------------------------------------------------------------------------

def _sym_direction(n, N, seed):
    R, _ = random_spd_cbtd(n, N, seed, off=1.0)
    D = O.cbtd_to_dense(R, n, N)
    return O.dense_to_cbtd(0.5 * (D + D.T), n, N)


def test_adjoint_matches_finite_differences():
    n, N = 3, 5
    S, D = random_spd_cbtd(n, N, 12)
    rng = np.random.default_rng(12)
    g = rng.normal(size=n * N)
    glam = rng.normal(size=n * N)
    lam = O.dense_solve(S, g, n, N)
    gg, gS = O.adjoint_dense(S, lam, glam, n, N)
    L = lambda S_, g_: float(glam @ O.dense_solve(S_, g_, n, N))
    eps = 1e-6
    for k in range(8):
        Dir = _sym_direction(n, N, 100 + k)
        fd = (L(S + eps * Dir, g) - L(S - eps * Dir, g)) / (2 * eps)
        assert np.isclose(fd, float(gS @ Dir), rtol=1e-6, atol=1e-9)
        dg = rng.normal(size=n * N)
        fdg = (L(S, g + eps * dg) - L(S, g - eps * dg)) / (2 * eps)
        assert np.isclose(fdg, float(gg @ dg), rtol=1e-6, atol=1e-9)
    # the pads of g_S are zero and the pattern is symmetric
    assert np.all(gS.reshape(N, 3, n * n)[0, 0] == 0) and np.all(gS.reshape(N, 3, n * n)[-1, 2] == 0)
    G = O.cbtd_to_dense(gS, n, N)
    assert np.allclose(G, G.T)


FIXTURES = ["lqr_n2_N3_jacobi.json", "lqr_n14_N32_ss.json", "lqr_n14_N128_ss.json"]


@pytest.mark.parametrize("fname", FIXTURES)
def test_oracle_reproduces_committed_fixtures(fname):
    """The committed fixtures (tests/golden/make_fixtures.py) hold the oracle's
    outputs on seeded LQR systems; the oracle must still produce them: FP64
    iteration counts and statuses exactly, lambda to 1e-12 and the r'z history to
    1e-9 relative; FP32 counts within one iteration (they depend on rounding)."""
    import json
    import gbdpcg
    d = json.load(open(os.path.join(HERE, "golden", fname)))
    gen = d["generator"]
    n, N = gen["n"], gen["N"]
    S, g = gbdpcg.lqr_generate(n, gen["m"], N, gen["seed"], dtype=np.float64)
    P = O.precond(S, n, N, gen["precond"])
    assert abs(np.linalg.norm(S) - d["checks"]["S_fro"]) <= 1e-12 * d["checks"]["S_fro"]
    assert abs(g.sum() - d["checks"]["gamma_sum"]) <= 1e-9 * max(1.0, abs(d["checks"]["gamma_sum"]))
    assert abs(np.linalg.norm(P) - d["checks"]["Pinv_fro"]) <= 1e-10 * d["checks"]["Pinv_fro"]
    for key, want in d["outputs"].items():
        mode_name, dt_name, form = key.split("_")
        dt = np.float64 if dt_name == "f64" else np.float32
        mode = O.EXIT_REL if mode_name == "rel" else O.EXIT_ABS
        lam, it, st, eta, rr, hist = O.pcg(S.astype(dt), P.astype(dt), g.astype(dt), np.zeros(n * N, dt), n, N,
                                           d["tol"], d["max_iter"], mode, dt, fused=form == "fused")
        if dt == np.float64:
            assert (it, st) == (want["iters"], want["status"]), (key, it, st)
        else:
            assert abs(it - want["iters"]) <= 1 and st == want["status"], (key, it, st)
        if "lambda" in want:
            ref = np.array(want["lambda"])
            assert np.linalg.norm(lam - ref) <= 1e-12 * np.linalg.norm(ref), key
            h = np.array(want["eta_history"])
            assert len(hist) == len(h) and np.allclose(hist, h, rtol=1e-9, atol=0), key
