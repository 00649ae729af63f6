"""The host-pointer drop-in (solvePCG(h_S, h_Pinv, h_gamma, h_lambda, n, N, cfg),
include/interface.cuh:24-91) on the device.

* Same bits as the device-pointer call (the inputs are staged in pinned host
  memory that the kernel reads in place; the kernel and layout are the same),
  with and without a Pinv argument, in FP32 and FP64.
* Reentrant on one stream: host threads calling it on the default stream (as
  every reference-style caller does) with different sizes and data each get the
  result of a solo call, bit for bit, also while the staging grows.
"""
import threading

import numpy as np
import pytest

from test_gpu_parity import gpu_solve, lqr_system

pytestmark = pytest.mark.gpu
KW = dict(tol=1e-5, max_iter=173, exit_mode=1, recurrence=1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("with_pinv", [True, False])
def test_host_pointer_equals_device_pointer(gpu, dtype, with_pinv):
    import gbdpcg
    n, N = 14, 128
    S, P, g = lqr_system(n, N, 2128, dtype)
    lam_d, res_d = gpu_solve(S, P if with_pinv else None, g, np.zeros(n * N, dtype), n, N, dtype, **KW)
    lam_h = np.zeros(n * N, dtype)
    S_before, g_before = S.copy(), g.copy()
    it = gbdpcg.solve_host(S, g, lam_h, n, N, P if with_pinv else None, **KW)
    assert it == res_d.iters and res_d.status == gbdpcg.CONVERGED
    assert np.array_equal(lam_h, lam_d)
    assert np.array_equal(S, S_before) and np.array_equal(g, g_before)  # inputs are read only


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_host_pointer_every_state_size(gpu, dtype):
    """The host-pointer call bit-identical to the device-pointer call for every
    stateSize 1..32 at N = 5 and 7 (staged arrays at offsets that are not
    16-byte multiples for most sizes; the staging grows along the way)."""
    import gbdpcg
    kw = dict(KW, tol=3e-5 if dtype == np.float32 else 1e-9, max_iter=400)
    for n in range(1, 33):
        for N in (5, 7):
            S, P, g = lqr_system(n, N, 5000 + 37 * n + N, dtype)
            lam_d, res_d = gpu_solve(S, P, g, np.zeros(n * N, dtype), n, N, dtype, **kw)
            lam_h = np.zeros(n * N, dtype)
            it = gbdpcg.solve_host(S, g, lam_h, n, N, P, **kw)
            assert res_d.status == gbdpcg.CONVERGED, (n, N)
            assert it == res_d.iters and np.array_equal(lam_h, lam_d), (n, N)


def test_host_pointer_reentrant_on_default_stream(gpu):
    """Three host threads, default stream, different (n, N) and data; one of them
    alternates sizes so the shared staging grows while the others run."""
    import gbdpcg
    dtype = np.float32
    shapes = [[(14, 64, 301)], [(6, 40, 302)], [(14, 32, 303), (12, 200, 304)]]
    solo = {}
    for group in shapes:
        for n, N, seed in group:
            S, P, g = lqr_system(n, N, seed, dtype)
            lam = np.zeros(n * N, dtype)
            it = gbdpcg.solve_host(S, g, lam, n, N, P, **KW)
            solo[(n, N, seed)] = (S, P, g, it, lam)
    errors, done = [], [0, 0, 0]

    def worker(i):
        try:
            for rep in range(24):
                n, N, seed = shapes[i][rep % len(shapes[i])]
                S, P, g, it_ref, lam_ref = solo[(n, N, seed)]
                lam = np.zeros(n * N, dtype)
                it = gbdpcg.solve_host(S, g, lam, n, N, P, **KW)
                if it != it_ref or not np.array_equal(lam, lam_ref):
                    errors.append((i, rep, n, N, it, it_ref))
                done[i] += 1
        except Exception as e:  # reported in the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not errors, errors[:5]
    assert done == [24, 24, 24]
