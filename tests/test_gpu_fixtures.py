"""Device solves held to the committed oracle fixtures (tests/golden/*.json,
written by tests/golden/make_fixtures.py and re-checked against the oracle in
tests/test_oracle.py): n=2 N=3 block-Jacobi, n=14 N=32 and N=128
symmetric-stair, relative-residual and absolute |r'z| rules, both recurrences,
FP32 and FP64.

Bar: the device count within +-1 of the fixture's count for the same form and
precision and of the reference form's count; in FP64, lambda within 1e-10 of
the fixture's lambda when the device ran the fixture's number of iterations.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import gpu_solve, rel

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = ["lqr_n2_N3_jacobi.json", "lqr_n14_N32_ss.json", "lqr_n14_N128_ss.json"]


@pytest.mark.parametrize("fname", FIXTURES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("mode", ["rel", "abs"])
@pytest.mark.parametrize("recurrence", [0, 1])
def test_device_against_fixture(gpu, fname, dtype, mode, recurrence):
    import gbdpcg
    d = json.load(open(os.path.join(HERE, "golden", fname)))
    gen = d["generator"]
    n, N = gen["n"], gen["N"]
    S, g = gbdpcg.lqr_generate(n, gen["m"], N, gen["seed"], dtype=np.float64)
    P = O.precond(S, n, N, gen["precond"])
    dt_name = "f64" if dtype == np.float64 else "f32"
    form = "fused" if recurrence == 1 else "reference"
    want = d["outputs"][f"{mode}_{dt_name}_{form}"]
    want_ref = d["outputs"][f"{mode}_{dt_name}_reference"]
    exit_mode = gbdpcg.EXIT_REL_RESID if mode == "rel" else gbdpcg.EXIT_ABS_ETA
    lam, res = gpu_solve(S.astype(dtype), P.astype(dtype), g.astype(dtype), np.zeros(n * N, dtype), n, N, dtype,
                         tol=d["tol"], max_iter=d["max_iter"], exit_mode=exit_mode, recurrence=recurrence)
    assert res.status == want["status"] == gbdpcg.CONVERGED
    assert abs(res.iters - want["iters"]) <= 1 and abs(res.iters - want_ref["iters"]) <= 1, \
        (res.iters, want["iters"], want_ref["iters"])
    fixture = d["outputs"][f"{mode}_f64_reference"]
    if dtype == np.float64 and res.iters == fixture["iters"]:
        ref = np.array(fixture["lambda"])
        assert rel(lam, ref) <= 1e-10
    else:  # the FP64 oracle run for the device's number of iterations
        lam_ref, *_ = O.pcg(S, P, g, np.zeros(n * N), n, N, 0.0, res.iters)
        assert rel(lam, lam_ref) <= (1e-4 if dtype == np.float32 else 1e-10)
