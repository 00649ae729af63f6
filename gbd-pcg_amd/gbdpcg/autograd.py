"""torch.autograd wrapper of the block-tridiagonal PCG solve.

lambda = S^-1 gamma with S in cbtd storage.  The backward pass launches the
library's adjoint solve (S g_gamma = g_lambda on the device, S and Pinv reused
on chip) and returns
  dL/dgamma = g_gamma,
  dL/dS     = -1/2 (g_gamma lambda^T + lambda g_gamma^T) on the block-tridiagonal
              pattern (cbtd layout, padding blocks zero),
which is the gradient for perturbations of S that stay symmetric and
block-tridiagonal.  Pinv gets no gradient (the exact solution does not depend on
it).  All work runs in the HIP library; tensors must be on the device.  A
forward or adjoint solve that does not converge raises RuntimeError.
"""
from __future__ import annotations

import torch

from . import CONVERGED, EXIT_REL_RESID, RECUR_FUSED, adjoint, solve

_STATUS = {0: "CONVERGED", 1: "MAX_ITER", 2: "BREAKDOWN", 3: "NONFINITE", 4: "TIMED_OUT"}


def _require_converged(what: str, status: int, iters: int) -> None:
    """An unconverged solve would hand back a wrong lambda or gradient without
    any sign of it: raise instead."""
    if status != CONVERGED:
        raise RuntimeError(f"PCGSolve: the {what} solve ended with status {_STATUS.get(status, status)} after "
                           f"{iters} iterations (raise max_iter or loosen tol)")


class PCGSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, S, gamma, Pinv, n, N, tol=1e-5, max_iter=173, exit_mode=EXIT_REL_RESID,
                recurrence=RECUR_FUSED):
        S_ = S.detach().contiguous()
        P_ = Pinv.detach().contiguous()
        lam = torch.zeros_like(gamma)
        res = solve(S_, gamma.detach().contiguous(), lam, n, N, P_, tol=tol, max_iter=max_iter,
                    exit_mode=exit_mode, recurrence=recurrence)
        _require_converged("forward", res.status, res.iters)
        ctx.save_for_backward(S_, P_, lam)
        ctx.cfg = (n, N, tol, max_iter, exit_mode, recurrence)
        ctx.forward_result = res
        return lam

    @staticmethod
    def backward(ctx, glam):
        S_, P_, lam = ctx.saved_tensors
        n, N, tol, max_iter, exit_mode, recurrence = ctx.cfg
        gg = torch.zeros_like(lam)
        gS = torch.zeros_like(S_)
        res = adjoint(S_, None, lam, glam.contiguous(), gg, gS, n, N, P_, forward=False, tol=tol,
                      max_iter=max_iter, exit_mode=exit_mode, recurrence=recurrence)
        _require_converged("adjoint", res.adj_status, res.adj_iters)
        return gS, gg, None, None, None, None, None, None, None


def pcg_solve(S, gamma, Pinv, n, N, **kw):
    """Differentiable solve: returns lambda with S^-1 gamma (see module doc)."""
    return PCGSolve.apply(S, gamma, Pinv, n, N, kw.get("tol", 1e-5), kw.get("max_iter", 173),
                          kw.get("exit_mode", EXIT_REL_RESID), kw.get("recurrence", RECUR_FUSED))
