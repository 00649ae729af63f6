"""Writes the committed oracle fixtures of this directory (run from the repository
root: python tests/golden/make_fixtures.py).

Each fixture is an LQR Schur-complement system from the library's seeded
generator (gbdpcg.lqr_generate), its preconditioner from the oracle's
definition (oracle.precond), and the oracle's outputs for the reference
iteration (include/pcg.cuh:118-208) and its fused form: iteration count,
status (FP64 and FP32, both forms) and, for the FP64 reference form, the final
lambda and the r'z history.  The systems' inputs
are identified by their generator arguments plus checksums (Frobenius norm of
S, sum of gamma); tests/test_oracle.py regenerates them and checks that the
oracle still reproduces these outputs, so a later change of the oracle cannot
drift silently, and tests/test_gpu_fixtures.py holds the device solves to them.

Fixtures:
  lqr_n2_N3_jacobi.json   n=2, m=1, N=3, seed 1000, block-Jacobi Pinv (the shape
                          of the reference's compile example, examples/pcg_solve.cu)
  lqr_n14_N32_ss.json     n=14, m=7, N=32, seed 2032, symmetric-stair Pinv
  lqr_n14_N128_ss.json    n=14, m=7, N=128, seed 2128, symmetric-stair Pinv
all at relative residual 1e-5 (max_iter 173) and at the reference's absolute
rule |r'z| < 1e-5.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "gbd-pcg_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

CASES = [
    ("lqr_n2_N3_jacobi.json", dict(n=2, m=1, N=3, seed=1000, precond=1)),
    ("lqr_n14_N32_ss.json", dict(n=14, m=7, N=32, seed=2032, precond=2)),
    ("lqr_n14_N128_ss.json", dict(n=14, m=7, N=128, seed=2128, precond=2)),
]
TOL, MAX_ITER = 1e-5, 173


def system(spec):
    import gbdpcg
    from oracle import oracle as O
    S, g = gbdpcg.lqr_generate(spec["n"], spec["m"], spec["N"], spec["seed"], dtype=np.float64)
    P = O.precond(S, spec["n"], spec["N"], spec["precond"])
    return S, P, g


def outputs(S, P, g, n, N):
    from oracle import oracle as O
    out = {}
    for mode_name, mode in (("rel", O.EXIT_REL), ("abs", O.EXIT_ABS)):
        for dt_name, dt in (("f64", np.float64), ("f32", np.float32)):
            for form, fused in (("reference", False), ("fused", True)):
                lam, it, st, eta, rr, hist = O.pcg(S.astype(dt), P.astype(dt), g.astype(dt), np.zeros(n * N, dt), n,
                                                   N, TOL, MAX_ITER, mode, dt, fused=fused)
                rec = {"iters": int(it), "status": int(st)}
                if dt == np.float64 and not fused:
                    rec["lambda"] = [float(v) for v in lam]
                    rec["eta_history"] = [float(v) for v in hist]
                out[f"{mode_name}_{dt_name}_{form}"] = rec
    return out


def main():
    for fname, spec in CASES:
        S, P, g = system(spec)
        n, N = spec["n"], spec["N"]
        data = {"generator": dict(spec, dt=0.1, dtype="float64"), "tol": TOL, "max_iter": MAX_ITER,
                "checks": {"S_fro": float(np.linalg.norm(S)), "gamma_sum": float(g.sum()),
                           "Pinv_fro": float(np.linalg.norm(P))},
                "outputs": outputs(S, P, g, n, N)}
        with open(os.path.join(HERE, fname), "w") as f:
            json.dump(data, f, indent=1)
            f.write("\n")
        o = data["outputs"]
        print(fname, {k: (v["iters"], v["status"]) for k, v in o.items()})


if __name__ == "__main__":
    main()
