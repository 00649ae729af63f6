"""The no-allocation solve (gbdpcg_solve_device_noalloc_*) through the C ABI,
against the stream-ordered solve it mirrors (gbdpcg_solve_device_async_*):
lambda and the result record bit-identical, for two calls in a row whose
second system is smaller (the shared scratch keeps the first call's records),
in FP32 and FP64; a system whose exchange records exceed the scratch is
rejected before any launch with lambda untouched; a missing Pinv is rejected.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).astype(dtype))).to(
        "cuda", torch.float64 if dtype == np.float64 else torch.float32)


def _system(n, N, seed, dtype):
    import gbdpcg
    S, g = gbdpcg.lqr_generate(n, max(n // 2, 1), N, seed, dtype=np.float64)
    return _dev(S, dtype), _dev(O.precond(S, n, N, 2), dtype), _dev(g, dtype)


def _record(t):
    import gbdpcg
    return gbdpcg.Result.from_buffer_copy(t.cpu().numpy().tobytes())


def _call(entry, n, N, S, P, g, lam, cfg, rec):
    return entry(n, N, S.data_ptr(), P.data_ptr(), g.data_ptr(), lam.data_ptr(), ctypes.byref(cfg),
                 ctypes.c_void_p(rec.data_ptr()) if rec is not None else None)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("recurrence", [0, 1])
def test_noalloc_matches_async_solve(gpu, dtype, recurrence):
    import torch
    import gbdpcg
    L = gbdpcg.lib()
    sfx = "f64" if dtype == np.float64 else "f32"
    f_async = getattr(L, f"gbdpcg_solve_device_async_{sfx}")
    f_noalloc = getattr(L, f"gbdpcg_solve_device_noalloc_{sfx}")
    stream = torch.cuda.current_stream()
    cfg = gbdpcg.make_config(tol=1e-5 if dtype == np.float32 else 1e-10, max_iter=300,
                             exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=recurrence,
                             stream=stream.cuda_stream or None)
    # the second call is smaller: fewer workgroups, fewer record words than the
    # first left in the shared scratch; the third is smaller in n as well
    for n, N, seed in ((14, 256, 11), (14, 64, 12), (6, 40, 13)):
        S, P, g = _system(n, N, seed, dtype)
        lam_a = torch.zeros(n * N, dtype=S.dtype, device="cuda")
        lam_b = torch.zeros(n * N, dtype=S.dtype, device="cuda")
        rec_a = torch.zeros(gbdpcg.RESULT_BYTES, dtype=torch.uint8, device="cuda")
        rec_b = torch.full((gbdpcg.RESULT_BYTES,), 0xAB, dtype=torch.uint8, device="cuda")
        assert _call(f_async, n, N, S, P, g, lam_a, cfg, rec_a) == 0, gbdpcg.last_error()
        assert _call(f_noalloc, n, N, S, P, g, lam_b, cfg, rec_b) == 0, gbdpcg.last_error()
        torch.cuda.synchronize()
        ra, rb = _record(rec_a), _record(rec_b)
        assert ra.status == gbdpcg.CONVERGED
        assert (rb.iters, rb.status, rb.eta, rb.rr) == (ra.iters, ra.status, ra.eta, ra.rr)
        assert torch.equal(lam_b, lam_a), (n, N)
        # without a record of its own the result goes to the scratch: lambda is the same
        lam_c = torch.zeros(n * N, dtype=S.dtype, device="cuda")
        assert _call(f_noalloc, n, N, S, P, g, lam_c, cfg, None) == 0, gbdpcg.last_error()
        torch.cuda.synchronize()
        assert torch.equal(lam_c, lam_a)


def test_noalloc_rejects_what_does_not_fit(gpu):
    import torch
    import gbdpcg
    # 512 workgroups of records, more than the 1 MiB scratch holds: the layout of
    # the 512-thread FP32 n = 14 register instance at two workgroups per CU (it
    # must stay within 128 VGPRs; with fewer co-resident workgroups the solve
    # falls to the LDS kernel, whose records fit)
    n, N = 14, 4096
    S, P, g = _system(n, N, 21, np.float32)
    lam = torch.full((n * N,), 2.0, device="cuda")
    rec = torch.zeros(gbdpcg.RESULT_BYTES, dtype=torch.uint8, device="cuda")
    cfg = gbdpcg.make_config(tol=1e-5, max_iter=173, exit_mode=gbdpcg.EXIT_REL_RESID,
                             recurrence=gbdpcg.RECUR_FUSED)
    rc = _call(gbdpcg.lib().gbdpcg_solve_device_noalloc_f32, n, N, S, P, g, lam, cfg, rec)
    assert rc == gbdpcg.ERR_TOO_LARGE and "scratch" in gbdpcg.last_error()
    torch.cuda.synchronize()
    assert bool((lam == 2.0).all())
    # a missing Pinv is rejected (forming it would need device memory)
    rc = gbdpcg.lib().gbdpcg_solve_device_noalloc_f32(n, N, S.data_ptr(), None, g.data_ptr(), lam.data_ptr(),
                                                      ctypes.byref(cfg), None)
    assert rc == gbdpcg.ERR_INVALID
    # and a fitting system right after works
    S2, P2, g2 = _system(14, 128, 22, np.float32)
    lam2 = torch.zeros(14 * 128, device="cuda")
    gbdpcg.solve_noalloc(S2, g2, lam2, 14, 128, P2, rec, tol=1e-5, max_iter=173, exit_mode=gbdpcg.EXIT_REL_RESID,
                         recurrence=gbdpcg.RECUR_FUSED)
    torch.cuda.synchronize()
    assert _record(rec).status == gbdpcg.CONVERGED


def test_noalloc_captured_as_first_call(gpu):
    """The entry's callers include code under HIP graph capture: captured as the
    first call for its shape (nothing may allocate inside capture), replayed
    three times with eager work in between, every replay equals the eager async
    solve bit for bit, result record included."""
    import torch
    import gbdpcg
    n, N = 14, 72  # a shape no other test solves, so nothing was planned for it before
    S, P, g = _system(n, N, 31, np.float32)
    kw = dict(tol=1e-5, max_iter=173, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    stream = torch.cuda.Stream()
    lam = torch.zeros(n * N, device="cuda")
    rec = torch.zeros(gbdpcg.RESULT_BYTES, dtype=torch.uint8, device="cuda")
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=stream):
        lam.zero_()
        gbdpcg.solve_noalloc(S, g, lam, n, N, P, rec, stream=stream, **kw)
    lam_ref = torch.zeros(n * N, device="cuda")
    rec_ref = torch.zeros(gbdpcg.RESULT_BYTES, dtype=torch.uint8, device="cuda")
    gbdpcg.solve_async(S, g, lam_ref, n, N, P, rec_ref, **kw)
    torch.cuda.synchronize()
    ref = _record(rec_ref)
    assert ref.status == gbdpcg.CONVERGED
    for _ in range(3):
        junk = torch.randn(1 << 16, device="cuda")
        junk.mul_(2.0)
        lam.fill_(7.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        r = _record(rec)
        assert (r.iters, r.status, r.eta, r.rr) == (ref.iters, ref.status, ref.eta, ref.rr)
        assert torch.equal(lam, lam_ref)
