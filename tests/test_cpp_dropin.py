"""The C++ drop-in facade (include/gpu_pcg.hpp) through the example program
gbd-pcg_amd/examples/pcg_solve.cpp, built by build().

CPU: without a device every call returns GBDPCG_ERR_DEVICE with a message and
the process exits normally (the reference would exit() inside the library).
GPU: the reference example (n=2, N=3) matches its closed form and the FP64
run takes 6 iterations; the documented pcg_solve with the device-built
symmetric-stair preconditioner matches the FP64 direct solve.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gbd-pcg_amd", "gbdpcg", "pcg_solve_example")


def run():
    if not os.path.exists(EXE):
        pytest.skip("example not built (run build())")
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout


def parse(out):
    res = {}
    for line in out.splitlines():
        name, rest = line.split(" ", 1)
        res[name] = rest
    return res


def test_facade_reports_errors_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    rc, out = run()
    assert rc == 1
    lines = out.strip().splitlines()
    assert len(lines) == 3 and all(" error -3: " in l for l in lines)


@pytest.mark.gpu
def test_facade_on_device(gpu):
    rc, out = run()
    assert rc == 0, out
    r = parse(out)
    d = json.load(open(os.path.join(ROOT, "tests", "golden", "readme_example.json")))
    exact = np.array(d["lambda_closed_form"])
    it64 = int(r["readme_fp64"].split()[0].split("=")[1])
    lam64 = np.array([float(v) for v in r["readme_fp64"].split("lambda=")[1].split()])
    lam32 = np.array([float(v) for v in r["readme_fp32"].split("lambda=")[1].split()])
    assert it64 == 6
    assert np.linalg.norm(lam64 - exact) / np.linalg.norm(exact) < 1e-8
    assert np.linalg.norm(lam32 - exact) / np.linalg.norm(exact) < 1e-4
    import gbdpcg
    S, g = gbdpcg.lqr_generate(14, 7, 128, 2128, dtype=np.float64)
    x = O.block_solve(S, g, 14, 128)
    f = r["lqr_n14_N128_fp64"].split()
    nrm = float(f[1].split("=")[1])
    assert abs(nrm - np.linalg.norm(x)) / np.linalg.norm(x) < 1e-9
