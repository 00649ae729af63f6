#!/usr/bin/env python3
"""Benchmark of the block-tridiagonal PCG solve on one MI355X.

Headline (value): microseconds per plain single solve of the LQR Schur system
n=14 (m=7), N=128, FP32, symmetric-stair preconditioner, exit at relative
residual 1e-5 (max_iter 173), lambda0 = 0, inputs resident in HBM, one solve per
step launched back to back through the drop-in device entry point.

Also reported on the same line:
  * roofline: HBM bytes of the solve's algorithmic traffic (S and Pinv read once,
    gamma read, lambda written) / the solve kernel's average duration (HIP events
    around each launch on the launch stream);
  * batched: forward + adjoint solves per second for 64 systems of the same shape
    in one persistent launch (graph-replayed), and its ratio to back-to-back
    plain solves;
  * cpu_baseline: the FP64 host restatement of the reference iteration (oracle/,
    single thread) on a time-bounded sample of the same workload.

  * uneven_512: 512 systems, n from {6, 12, 14, 28}, N uniform in [16, 256],
    forward solves at tol 1e-5, the autotuned plan (one persistent launch or
    one per kernel class) graph-replayed; the untuned single launch beside it;
  * mpc_sweep_n28_N512: 256 systems n=28, N=512, forward + adjoint, ten
    warm-started re-solves (lambda shifted one knot, gamma += 1e-2 N(0,1)) in one
    graph, FP32 and FP64.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-batched] [--no-workloads]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gbd-pcg_amd"))

METRIC = "µs/solve to 1e-5 rel. resid (n=14,N=128); batched fwd+adjoint solves/sec"
HBM_PEAK_GBS = 8000.0
# LDS: 256 CUs x 256 B/clk (gfx950's LDS read rate; writes are half that) x 2.4 GHz
LDS_PEAK_GBS = 256 * 256 * 2.4
# FP32 VALU without packed-FP32 instructions (the build disables them): 256 CUs x 64
# lanes x 2 flops (FMA) x 2.4 GHz
VALU_F32_PEAK_TFLOPS = 256 * 64 * 2 * 2.4e9 / 1e12


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-workloads", action="store_true", help="skip the 512-system and n=28/N=512 workloads")
    ap.add_argument("--no-tune", action="store_true", help="time the planner's default layout, not the tuned one")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the other single-solve sizes (N = 32, 64, 256), e.g. for per-kernel profiles")
    ap.add_argument("--no-load-phase", action="store_true",
                    help="skip the max_iter = 0 launches (set-up and load/store phases alone), e.g. so that a "
                         "kernel trace holds full solves only")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the synchronous and host-pointer drop-in legs (their launches would mix into a "
                         "kernel trace of the device-pointer solves; the host-pointer call's kernel reads its "
                         "inputs over PCIe)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the solver runs each system on one "
              "device (a batch would shard by systems, one plan per GPU, with no exchange between them).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import gbdpcg

    n, m, N, tol, max_iter = 14, 7, 128, 1e-5, 173
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

    # synthetic LQR Schur system (seeded), SS preconditioner formed on the device
    S_h, g_h = gbdpcg.lqr_generate(n, m, N, 2128, dtype=np.float32)
    S = torch.from_numpy(S_h).to(dev)
    g = torch.from_numpy(g_h).to(dev)
    P = torch.empty_like(S)
    gbdpcg.precond_device(S, P, n, N, gbdpcg.PRECOND_SS)
    kw = dict(tol=tol, max_iter=max_iter, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)

    # on-device symmetric-stair Pinv formation (the documented call without Pinv pays it)
    Ptmp = torch.empty_like(S)
    for _ in range(3):
        gbdpcg.precond_device(S, Ptmp, n, N, gbdpcg.PRECOND_SS)
    pa, pb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    pa.record(stream)
    for _ in range(20):
        gbdpcg.precond_device(S, Ptmp, n, N, gbdpcg.PRECOND_SS)
    pb.record(stream)
    torch.cuda.synchronize()
    precond_us = pa.elapsed_time(pb) * 1000.0 / 20
    del Ptmp

    # the autotuner picks the layout of this shape (timed on these inputs)
    tune = None if args.no_tune else gbdpcg.tune_single(S, g, n, N, P, **kw)
    # latency floor of one exchange round at this layout (exchange alone, no work)
    floor = gbdpcg.exchange_floor(tune["wgs"] if tune else 16, tune["threads"] if tune else 256, rounds=4000)
    # the device's rates measured on streaming kernels (FP32 FMA, LDS reads, HBM reads)
    rates = gbdpcg.measure_rates()

    # one zeroed lambda per step so every step is a full solve from lambda0 = 0
    K, W = args.steps, args.warmup
    lams = torch.zeros((K + W, n * N), device=dev)
    dres = torch.zeros(gbdpcg.RESULT_BYTES, dtype=torch.uint8, device=dev)
    res = gbdpcg.solve(S, g, torch.zeros(n * N, device=dev), n, N, P, **kw)
    iters = res.iters
    torch.cuda.synchronize()
    for i in range(W):
        gbdpcg.solve_async(S, g, lams[i], n, N, P, dres, **kw)
    torch.cuda.synchronize()
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for i in range(K):
        gbdpcg.solve_async(S, g, lams[W + i], n, N, P, dres, **kw)
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    us_per_solve = total_ms * 1000.0 / K
    lam_last = lams[W + K - 1].cpu().numpy()
    # the solve kernel's own duration: HIP events around each launch on the launch
    # stream (a second pass; the events would otherwise count in the steps above)
    kn = min(K, 200)
    lams[:kn].zero_()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * kn)]
    torch.cuda.synchronize()
    for i in range(kn):
        ev[2 * i].record(stream)
        gbdpcg.solve_async(S, g, lams[i], n, N, P, dres, **kw)
        ev[2 * i + 1].record(stream)
    torch.cuda.synchronize()
    kern_each = [ev[2 * i].elapsed_time(ev[2 * i + 1]) * 1000.0 for i in range(kn)]
    kern_us = float(np.mean(kern_each))

    # the load phase alone (max_iter = 0: S and Pinv staged from HBM into registers,
    # r0 formed, one exchange, lambda written back), timed the same way
    load_us = None
    if not args.no_load_phase:
        kw0 = dict(kw, max_iter=0)
        ev0 = [torch.cuda.Event(enable_timing=True) for _ in range(2 * kn)]
        torch.cuda.synchronize()
        for i in range(kn):
            ev0[2 * i].record(stream)
            gbdpcg.solve_async(S, g, lams[i], n, N, P, dres, **kw0)
            ev0[2 * i + 1].record(stream)
        torch.cuda.synchronize()
        load_us = float(np.median([ev0[2 * i].elapsed_time(ev0[2 * i + 1]) * 1000.0 for i in range(kn)]))
        lams[:kn].zero_()
    # one iteration: the kernel less its fixed part (the max_iter = 0 launch)
    per_iter_us = (float(np.median(kern_each)) - (load_us or 0.0)) / max(iters, 1)

    # end-to-end drop-in calls through the C ABI with their arguments prepared
    # beforehand (the binding's per-call argument checks are Python overhead a
    # C or C++ caller does not pay): the synchronous device-pointer call (launch,
    # wait, result record) and the host-pointer call (inputs staged in, lambda
    # out, PCIe included), each from lambda0 = 0
    import ctypes
    P_h = P.cpu().numpy()
    sync_us = host_us = None
    sync_each = host_each = [0.0]
    if not args.no_dropin:
        sync_us, sync_each, host_us, host_each = dropin_legs(gbdpcg, torch, np, ctypes, S, P, g, S_h, g_h, n, N, kw,
                                                             stream, iters, dev)
    # the reference recurrence (two reductions per iteration) through the same entry point
    kw_ref = dict(kw, recurrence=gbdpcg.RECUR_REFERENCE)
    nref = min(K, 200)
    lams.zero_()
    for i in range(W):
        gbdpcg.solve_async(S, g, lams[i], n, N, P, dres, **kw_ref)
    torch.cuda.synchronize()
    t_start.record(stream)
    for i in range(nref):
        gbdpcg.solve_async(S, g, lams[W + i], n, N, P, dres, **kw_ref)
    t_end.record(stream)
    torch.cuda.synchronize()
    ref_us = t_start.elapsed_time(t_end) * 1000.0 / nref
    lams[:W + nref].zero_()

    # the other single-solve sizes of the same problem family (n=14, m=7), tuned,
    # back to back from lambda0 = 0
    sweep = {}
    for Ns in () if args.no_sweep else (32, 64, 128, 256):
        Ss_h, gs_h = gbdpcg.lqr_generate(n, m, Ns, 2000 + Ns, dtype=np.float32)
        Ss, gs = torch.from_numpy(Ss_h).to(dev), torch.from_numpy(gs_h).to(dev)
        Ps = torch.empty_like(Ss)
        gbdpcg.precond_device(Ss, Ps, n, Ns, gbdpcg.PRECOND_SS)
        if not args.no_tune:
            gbdpcg.tune_single(Ss, gs, n, Ns, Ps, **kw)
        ls = torch.zeros((60, n * Ns), device=dev)
        for i in range(10):
            gbdpcg.solve_async(Ss, gs, ls[i], n, Ns, Ps, dres, **kw)
        torch.cuda.synchronize()
        t_start.record(stream)
        for i in range(10, 60):
            gbdpcg.solve_async(Ss, gs, ls[i], n, Ns, Ps, dres, **kw)
        t_end.record(stream)
        torch.cuda.synchronize()
        rs = gbdpcg.solve(Ss, gs, torch.zeros(n * Ns, device=dev), n, Ns, Ps, **kw)
        sweep[f"N{Ns}"] = {"us_per_solve": round(t_start.elapsed_time(t_end) * 1000.0 / 50, 2), "iters": rs.iters}

    # roofline of the solve kernel: algorithmic HBM bytes per launch (S and Pinv read
    # once, gamma and lambda0 read, lambda written)
    ts = 4
    hbm_bytes = (2 * 3 * n * n * N + 3 * n * N) * ts
    # the kernel's average launch duration: HIP events over the timed region of
    # back-to-back launches on the launch stream, over the launches (per-launch
    # event pairs, kernel_us, add the events' own overhead)
    achieved = hbm_bytes / (us_per_solve * 1e-6) / 1e9
    # measured HBM bytes per launch: rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this
    # same command, scaled by the counters' reading of a plain copy of known size
    # (the counters are looked up under the name of the kernel this layout runs)
    layout = {"kernel": tune["kernel"] if tune else gbdpcg.KERNEL_REGISTER,
              "threads": tune["threads"] if tune else 256, "wgs": tune["wgs"] if tune else None}
    kname = {gbdpcg.KERNEL_LDS: "gbdpcg::pcg_kernel<float>",
             gbdpcg.KERNEL_REGISTER: f"gbdpcg::pcg_reg_kernel<float, 14, {layout['threads']}, true>",
             gbdpcg.KERNEL_ROWCAST: f"gbdpcg::pcg_rowcast_kernel<float, 14, {layout['threads']}, true>"}[
        layout["kernel"]]
    layout["kernel_name"] = kname
    traffic, traffic_kernel, load_traffic = None, None, None
    counters = os.path.join(ROOT, "profiles", "pcg_reg_kernel_n14_N128_counters.json")
    if os.path.exists(counters):
        with open(counters) as f:
            kernels = json.load(f)["kernels"]
        # the same kernel at the other thread count when this layout was not the
        # profiled one (the tuner's choice between them is within noise)
        family = kname.split("<")[0] + "<float, 14, "
        for name in [kname] + sorted(k for k in kernels if k.startswith(family) and k.endswith(", true>")):
            if name in kernels:
                traffic, traffic_kernel = int(kernels[name]["traffic_bytes"]), name
                if kernels[name].get("load_phase_traffic_bytes"):
                    load_traffic = int(kernels[name]["load_phase_traffic_bytes"])
                break
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "traffic_kernel": traffic_kernel,
                "algorithmic_bytes": hbm_bytes, "hbm_read_measured": round(rates["hbm_read_gbs"], 1),
                "load_phase": None if load_us is None else {
                    "us": round(load_us, 3), "achieved": round(hbm_bytes / (load_us * 1e-6) / 1e9, 3),
                    "frac": round(hbm_bytes / (load_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 6),
                    "traffic": load_traffic,
                    "achieved_counters": None if load_traffic is None else
                    round(load_traffic / (load_us * 1e-6) / 1e9, 3),
                    "note": "max_iter = 0: matrices to registers, r0, two exchanges, lambda out; latency-bound at "
                            "this size (under 40 KB per workgroup); traffic = rocprofv3 FETCH_SIZE + WRITE_SIZE "
                            "of the max_iter = 0 dispatches (calibrated), achieved_counters = traffic / us"},
                "latency": {"per_iteration_us": round(per_iter_us, 3),
                            "exchange_floor_us": round(floor["us_per_round"], 3),
                            "frac": round(floor["us_per_round"] / per_iter_us, 4)},
                "note": "latency-bound: one cross-workgroup exchange per iteration; per_iteration_us (kernel "
                        "less its max_iter = 0 launch, over the iterations) against exchange_floor_us (the "
                        "exchange alone at the same layout)"}

    # batched forward + adjoint (64 systems, same shape), one persistent launch
    batched = None
    if not args.no_batched:
        B = 64
        systems, host64 = [], []
        for b in range(B):
            Sb_h, gb_h = gbdpcg.lqr_generate(n, m, N, 4000 + b, dtype=np.float32)
            Sb = torch.from_numpy(Sb_h).to(dev)
            Pb = torch.empty_like(Sb)
            gbdpcg.precond_device(Sb, Pb, n, N, gbdpcg.PRECOND_SS)
            glam_h = gbdpcg.randn(n * N, 7000 + b)
            systems.append(dict(S=Sb, Pinv=Pb, gamma=torch.from_numpy(gb_h).to(dev),
                                lam=torch.zeros(n * N, device=dev), n=n, N=N, tol=tol, max_iter=max_iter,
                                glam=torch.from_numpy(glam_h).to(dev),
                                ggamma=torch.zeros(n * N, device=dev), gS=torch.zeros_like(Sb), adjoint=True))
            host64.append((Sb_h.astype(np.float64), Pb.cpu().numpy().astype(np.float64), gb_h.astype(np.float64),
                           glam_h.astype(np.float64)))
        plan = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED,
                            autotune=not args.no_tune)
        plan.run()
        bres = plan.results()
        graph = torch.cuda.CUDAGraph()
        gs = torch.cuda.Stream()
        torch.cuda.synchronize()
        lams64 = [d["lam"] for d in systems]
        with torch.cuda.graph(graph, stream=gs):
            torch._foreach_zero_(lams64)  # one launch resets every lambda0
            plan.run(stream=gs)
        for _ in range(3):
            graph.replay()
        torch.cuda.synchronize()
        nb = 30
        t_start.record(stream)
        for _ in range(nb):
            graph.replay()
        t_end.record(stream)
        torch.cuda.synchronize()
        batch_us = t_start.elapsed_time(t_end) * 1000.0 / nb
        rate = B / (batch_us * 1e-6)
        # the batch kernel's own duration: events on the plan's stream around each
        # eager run (lambda0 zeroed before the first event)
        bev = []
        for _ in range(20):
            with torch.cuda.stream(gs):
                torch._foreach_zero_(lams64)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(gs)
            plan.run(stream=gs)
            b.record(gs)
            bev.append((a, b))
        torch.cuda.synchronize()
        bkern_each = [a.elapsed_time(b) * 1000.0 for a, b in bev]
        bres = plan.results()
        # accuracy of the timed plan's own outputs (the last eager run above), before
        # anything else touches them: lambda, g_gamma and g_S against FP64 Cholesky
        # solves and the exact gradient formed from them
        e_lam = e_gg = e_gS = 0.0
        for d, (Sb64, _, gb64, gl64) in zip(systems, host64):
            x = banded_cholesky_solve(np, Sb64, gb64, n, N)
            y = banded_cholesky_solve(np, Sb64, gl64, n, N)
            e_lam = max(e_lam, rel_err(np, d["lam"].cpu().numpy(), x))
            e_gg = max(e_gg, rel_err(np, d["ggamma"].cpu().numpy(), y))
            e_gS = max(e_gS, rel_err(np, d["gS"].cpu().numpy(), cbtd_grad_S(np, y, x, n, N)))
        # the load and store phases alone: the same plan at max_iter = 0 (S and Pinv
        # staged, r0 of both sides, one exchange, lambda, g_gamma and g_S written),
        # on output buffers of its own
        lsp = None
        if not args.no_load_phase:
            sys0 = []
            for d, w in zip(systems, plan.wgs()):  # the same workgroups per system as the timed plan
                sys0.append(dict(d, lam=torch.zeros_like(d["lam"]), ggamma=torch.zeros_like(d["ggamma"]),
                                 gS=torch.zeros_like(d["gS"]), max_iter=0, wgs=w))
            plan0 = gbdpcg.Batch(sys0, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED,
                                 threads=plan.tune_report["threads"] if plan.tune_report else 0)
            plan0_info = plan0.info()
            lev = []
            for _ in range(20):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(gs)
                plan0.run(stream=gs)
                b.record(gs)
                lev.append((a, b))
            torch.cuda.synchronize()
            bload_us = float(np.median([a.elapsed_time(b) * 1000.0 for a, b in lev]))
            plan0.close()
            del sys0
            # algorithmic HBM bytes of forward + adjoint per system: S, Pinv, gamma,
            # lambda0, g_lambda read; lambda, g_gamma, g_S written
            bbytes = B * (9 * n * n * N + 5 * n * N) * 4
            gbs = bbytes / (bload_us * 1e-6) / 1e9
            pair_name = f"gbdpcg::pcg_pair_kernel<float, 14, {(plan.tune_report or {}).get('threads') or 512}, false>"
            ltr = committed_counters().get(pair_name, {}).get("load_phase_traffic_bytes")
            lsp = {"us": round(bload_us, 2), "algorithmic_bytes": bbytes, "bound": "hbm", "achieved": round(gbs, 1),
                   "unit": "GB/s", "peak": HBM_PEAK_GBS, "frac": round(gbs / HBM_PEAK_GBS, 4),
                   "frac_of_measured": round(gbs / rates["hbm_read_gbs"], 4),
                   "traffic": None if ltr is None else int(ltr), "traffic_kernel": pair_name,
                   "achieved_counters": None if ltr is None else round(ltr / (bload_us * 1e-6) / 1e9, 1),
                   "plan": plan0_info}
        broof = batched_roofline(np, gbdpcg, plan, bres, n, N, float(np.median(bkern_each)), rates)
        batched = {"systems": B, "us_per_batch": round(batch_us, 2), "fwd_adj_solves_per_s": round(rate, 1),
                   "vs_back_to_back_plain": round(rate / (1e6 / us_per_solve), 2),
                   "mean_iters_fwd": float(np.mean([r.iters for r in bres])),
                   "mean_iters_adj": float(np.mean([r.adj_iters for r in bres])),
                   "all_converged": all(r.status == 0 and r.adj_status == 0 for r in bres),
                   "max_rel_err_vs_fp64_cholesky": {"lambda": e_lam, "g_gamma": e_gg, "g_S": e_gS,
                                                     "systems_checked": B},
                   "kernel_us": spread(np, bkern_each), "roofline": broof,
                   "load_store_phase": lsp,
                   "plan": plan.info(), "tune": plan.tune_report,
                   "cpu_baseline": cpu_batched(np, host64, n, N, tol, max_iter, args.cpu_seconds)}

    workloads = None
    if not args.no_workloads:
        wsec = args.cpu_seconds / 2
        workloads = {"lqr_n2_N3_jacobi": lqr_n2_N3(gbdpcg, torch, np, dev, wsec),
                     "uneven_512": uneven_512(gbdpcg, torch, np, dev, not args.no_tune, rates, wsec)}
        for dt in (np.float32, np.float64):
            workloads[f"mpc_sweep_n28_N512_{np.dtype(dt).name}"] = mpc_sweep(gbdpcg, torch, np, dev, dt, rates, wsec)

    if args.calibrate_copy:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch give the
        # counters' scale (256 MiB read, 256 MiB written)
        src = torch.ones(64 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # CPU baseline leg: the FP64 host restatement of the reference iteration
    # (oracle/), one thread, which also checks the last timed solve
    cpu, err = cpu_single(np, S_h, g_h, P_h, n, N, tol, max_iter, iters, lam_last, args.cpu_seconds)
    # error against an FP64 dense Cholesky solve of the same system (numpy / scipy)
    chol_err = dense_cholesky_error(np, S_h, g_h, n, N, lam_last)

    out = {
        "metric": METRIC, "value": round(us_per_solve, 3), "unit": "us/solve", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": round(us_per_solve / 1000.0, 6), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded LQR Schur complement, on-device SS Pinv)",
        "config": {"workload": "plain PCG single solve, n=14 m=7 N=128, symmetric-stair Pinv, rel resid 1e-5, "
                               "max_iter 173, lambda0=0", "n": n, "N": N, "recurrence": "fused (Chronopoulos-Gear)",
                   "iters": iters, "rel_err_vs_fp64_oracle": err, "rel_err_vs_fp64_dense_cholesky": chol_err},
        "roofline": roofline, "cpu_baseline": cpu,
        "kernel_us": round(kern_us, 3), "kernel_us_spread": spread(np, kern_each),
        "per_iteration_us": round(per_iter_us, 3),
        "exchange_floor_us": round(floor["us_per_round"], 3), "exchange_xcd_local": floor["xcd_local"],
        "sync_dropin_us": None if sync_us is None else round(sync_us, 2),
        "sync_dropin_us_spread": None if sync_us is None else spread(np, sync_each),
        "host_pointer_dropin_us": None if host_us is None else round(host_us, 2),
        "host_pointer_dropin_us_spread": None if host_us is None else spread(np, host_each),
        "reference_recurrence_us": round(ref_us, 2), "precond_ss_us": round(precond_us, 2),
        "measured_rates": {k: round(v, 2) for k, v in rates.items()}, "n14_sizes": sweep, "tune": tune, "batched": batched, "workloads": workloads,
    }
    failed = accuracy_failures(out)
    out["ok"] = not failed
    out["failed_checks"] = failed
    print(json.dumps(out), flush=True)
    if failed:
        print("bench.py: accuracy checks failed: " + "; ".join(failed), file=sys.stderr, flush=True)
        sys.exit(1)


# relative error bound of every reported result against its FP64 reference (the
# solves stop at a 1e-5 relative residual)
ERR_BOUND = 1e-4


def accuracy_failures(out):
    """Names of the checks the reported results fail: a solve that did not
    converge, or an error against the FP64 references above ERR_BOUND."""
    bad = []
    cfg = out["config"]
    for k in ("rel_err_vs_fp64_oracle", "rel_err_vs_fp64_dense_cholesky"):
        if not cfg[k] <= ERR_BOUND:
            bad.append(f"single solve {k} = {cfg[k]}")
    legs = {}
    if out.get("batched"):
        legs["batched"] = out["batched"]
    for name, w in (out.get("workloads") or {}).items():
        legs[name] = w
    for name, leg in legs.items():
        if "all_converged" in leg and not leg["all_converged"]:
            bad.append(f"{name}: not every system converged")
        for k, v in (leg.get("max_rel_err_vs_fp64_cholesky") or {}).items():
            if k != "systems_checked" and not v <= ERR_BOUND:
                bad.append(f"{name}: {k} rel err {v}")
        if leg.get("replay_equals_eager") is False:
            bad.append(f"{name}: graph replay differs from the eager sequence")
        alt = leg.get("one_launch_untuned")
        if alt and not alt.get("all_converged", True):
            bad.append(f"{name}: untuned single launch: not every system converged")
    return bad


def dropin_legs(gbdpcg, torch, np, ctypes, S, P, g, S_h, g_h, n, N, kw, stream, iters, dev):
    """The synchronous device-pointer call (launch, wait, result record) and the
    host-pointer call (inputs staged in, lambda out, PCIe included), each from
    lambda0 = 0, through the C ABI with their arguments prepared beforehand."""
    cfg_c = gbdpcg.make_config(stream=stream.cuda_stream or None, **kw)

    rec_c = gbdpcg.Result()
    reps = 100
    lam_rows = torch.zeros((reps + 10, n * N), device=dev)
    f_sync = gbdpcg.lib().gbdpcg_solve_device_f32
    args_c = [(n, N, S.data_ptr(), P.data_ptr(), g.data_ptr(), lam_rows[i].data_ptr(), ctypes.byref(cfg_c),
               ctypes.byref(rec_c)) for i in range(reps + 10)]
    for i in range(10):
        f_sync(*args_c[reps + i])
    torch.cuda.synchronize()
    sync_each = []
    for i in range(reps):
        t0 = time.perf_counter()
        rc = f_sync(*args_c[i])
        sync_each.append((time.perf_counter() - t0) * 1e6)
        assert rc == iters, (rc, gbdpcg.last_error())
    sync_us = float(np.mean(sync_each))
    P_h = P.cpu().numpy()
    cfg_h = gbdpcg.make_config(**kw)  # the reference's call shape: default stream
    f_host = gbdpcg.lib().gbdpcg_solve_f32
    lam_h = [np.zeros(n * N, dtype=np.float32) for _ in range(reps + 10)]
    host_each = []
    for i in range(reps + 10):
        t0 = time.perf_counter()
        rc = f_host(S_h.ctypes.data, P_h.ctypes.data, g_h.ctypes.data, lam_h[i].ctypes.data, n, N,
                    ctypes.byref(cfg_h))
        if i >= 10:
            host_each.append((time.perf_counter() - t0) * 1e6)
        assert rc == iters, (rc, gbdpcg.last_error())
    host_us = float(np.mean(host_each))
    assert np.array_equal(lam_h[-1], lam_rows[0].cpu().numpy())  # the same bits as the device-pointer call
    return sync_us, sync_each, host_us, host_each


def committed_counters():
    """Per-kernel figures of the committed rocprofv3 counter summary (empty when
    it is absent)."""
    path = os.path.join(ROOT, "profiles", "pcg_reg_kernel_n14_N128_counters.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("kernels", {})


def host_cpu():
    """The host's CPU model and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": usable}


def spread(np, xs):
    """Mean and 10th / 50th / 90th percentiles of per-launch times (us)."""
    p10, p50, p90 = np.percentile(np.asarray(xs, dtype=np.float64), [10, 50, 90])
    return {"mean": round(float(np.mean(xs)), 3), "p10": round(float(p10), 3), "p50": round(float(p50), 3),
            "p90": round(float(p90), 3), "launches": len(xs)}


def batched_roofline(np, gbdpcg, plan, res, n, N, kern_us, rates):
    """Roofline of the batched forward + adjoint launch.  The systems' vectors and
    blocks live in registers and LDS for the whole solve, so the bounds are the
    LDS rate and the FP32 VALU rate, not HBM:
      - LDS bytes per launch: rocprofv3 SQ_INSTS_LDS_LOAD_BANDWIDTH +
        SQ_INSTS_LDS_STORE_BANDWIDTH (64-byte units) of this kernel, from the
        committed counter summary, over the kernel time measured here;
      - algorithmic flops: per iteration and system, two block-tridiagonal
        products (S p and Pinv r: 2 x 2 x (3N - 2) n^2) and ten vector flops per
        entry (2 dot products, 3 axpys) = 4 (3N - 2) n^2 + 10 nN, times the
        forward plus adjoint iterations each system ran."""
    t = plan.tune_report or {}
    threads = t.get("threads") or 512
    fam = "gbdpcg::pcg_pair_kernel<float, 14, "
    name = f"{fam}{threads}, false>"
    lds, lds_kernel = None, None
    counters = os.path.join(ROOT, "profiles", "pcg_reg_kernel_n14_N128_counters.json")
    if os.path.exists(counters):
        with open(counters) as f:
            kernels = json.load(f)["kernels"]
        for k in [name] + sorted(k for k in kernels if k.startswith(fam)):
            if k in kernels and kernels[k].get("lds_bytes"):
                lds, lds_kernel = float(kernels[k]["lds_bytes"]), k
                break
    per_iter = 4 * (3 * N - 2) * n * n + 10 * n * N
    flops = per_iter * sum(r.iters + r.adj_iters for r in res)
    tflops = flops / (kern_us * 1e-6) / 1e12
    # S and Pinv are register-resident in this kernel: the VALU rate is its roofline
    out = {"bound": "valu", "achieved": round(tflops, 3), "peak": round(VALU_F32_PEAK_TFLOPS, 2),
           "unit": "TFLOP/s", "frac": round(tflops / VALU_F32_PEAK_TFLOPS, 4), "algorithmic_flops": flops,
           "peak_measured": round(rates["valu_f32_tflops"], 2),
           "frac_of_measured": round(tflops / rates["valu_f32_tflops"], 4),
           "lds": {"bytes": lds, "kernel": lds_kernel, "achieved": None, "peak": round(LDS_PEAK_GBS, 1),
                   "peak_measured": round(rates["lds_read_gbs"], 1), "unit": "GB/s", "frac": None},
           "note": "latency-bound like the single solve: one exchange round per iteration; these fractions "
                   "show how far the per-iteration work is from either resource"}
    if lds is not None:
        gbs = lds / (kern_us * 1e-6) / 1e9
        out["lds"]["achieved"], out["lds"]["frac"] = round(gbs, 1), round(gbs / LDS_PEAK_GBS, 4)
    return out


def cpu_single(np, S_h, g_h, P_h, n, N, tol, max_iter, iters, lam_last, seconds):
    """CPU baseline of the single solve: the FP64 host restatement of the
    reference iteration (oracle/), one thread, timed on a bounded sample; as the
    checker it also gives the last GPU solve's error against the FP64 iteration
    run for the same number of iterations."""
    from oracle import oracle as O
    S64, g64, P64 = S_h.astype(np.float64), g_h.astype(np.float64), P_h.astype(np.float64)
    lam0 = np.zeros(n * N)
    for _ in range(10):  # warm-up
        O.pcg(S64, P64, g64, lam0, n, N, tol, max_iter, O.EXIT_REL)
    cnt, each, t0 = 0, [], time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        t1 = time.perf_counter()
        _, it_cpu, *_ = O.pcg(S64, P64, g64, lam0, n, N, tol, max_iter, O.EXIT_REL)
        each.append((time.perf_counter() - t1) * 1e6)
        cnt += 1
    cpu_us = (time.perf_counter() - t0) * 1e6 / cnt
    cpu = {"value": round(cpu_us, 2), "unit": "us/solve", "cores": 1, "kind": "port",
           "sample": f"{cnt} FP64 solves of the same n=14,N=128 system (oracle/pcg_oracle.c, {it_cpu} iterations each)",
           "spread": spread(np, each), "host": host_cpu()}
    lam_ref, *_ = O.pcg(S64, P64, g64, lam0, n, N, 0.0, iters)
    err = float(np.linalg.norm(lam_last - lam_ref) / np.linalg.norm(lam_ref))
    return cpu, err


def banded_cholesky_solve(np, S_h, rhs, n, N):
    """FP64 Cholesky solve of S x = rhs in banded storage (half-bandwidth 2n - 1:
    the same factorisation as the dense one, restricted to the band), for the
    larger configurations."""
    from scipy.linalg import solveh_banded
    B = np.asarray(S_h, dtype=np.float64).reshape(N, 3, n, n).transpose(0, 1, 3, 2)  # [k, block, row, column]
    u = 2 * n - 1
    ab = np.zeros((u + 1, n * N))
    r, c = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    k = np.arange(N)[:, None]
    up = r <= c  # diagonal blocks: their upper triangle
    ab[(u + r[up] - c[up])[None, :], k * n + c[up][None, :]] = B[:, 1][:, up]
    # blocks (k, k+1), k < N-1: wholly above the diagonal
    if N > 1:
        ab[(u + r - c - n).ravel()[None, :], (k[:-1] + 1) * n + c.ravel()[None, :]] = B[:-1, 2].reshape(N - 1, -1)
    return solveh_banded(ab, np.asarray(rhs, dtype=np.float64))


def rel_err(np, a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def cbtd_grad_S(np, gg, lam, n, N):
    """-1/2 (gg lam^T + lam gg^T) on the block-tridiagonal pattern, in cbtd
    storage (block row k: blocks (k, k-1), (k, k), (k, k+1), each n x n
    column-major; the two padding blocks zero)."""
    G = np.asarray(gg, np.float64).reshape(N, n)
    L = np.asarray(lam, np.float64).reshape(N, n)
    out = np.zeros((N, 3, n, n))  # [k, block, column, row]
    for b, off in enumerate((-1, 0, 1)):
        k = np.arange(max(0, -off), N - max(0, off))
        j = k + off
        # element (row r of knot k, column c of knot j) sits at [k, b, c, r]
        out[k, b] = -0.5 * (G[k][:, None, :] * L[j][:, :, None] + L[k][:, None, :] * G[j][:, :, None])
    return out.reshape(-1)


def dense_cholesky_error(np, S_h, g_h, n, N, lam):
    """Relative error of lam against the FP64 dense Cholesky solution of S x = gamma
    (S expanded from its cbtd blocks: block row k holds blocks (k, k-1), (k, k),
    (k, k+1), each n x n column-major)."""
    from scipy.linalg import cho_factor, cho_solve
    B = S_h.astype(np.float64).reshape(N, 3, n, n)  # [k, block, column, row]
    D = np.zeros((n * N, n * N))
    for k in range(N):
        for b in range(3):
            j = k - 1 + b
            if 0 <= j < N:
                D[k * n:(k + 1) * n, j * n:(j + 1) * n] = B[k, b].T
    x = cho_solve(cho_factor(D), g_h.astype(np.float64))
    return float(np.linalg.norm(lam.astype(np.float64) - x) / np.linalg.norm(x))


def cpu_batched(np, host, n, N, tol, max_iter, seconds):
    """The same 64 forward + adjoint solves on the host: the FP64 restatement of
    the reference iteration (oracle/), one system per thread on the host's cores
    (the C solver releases the interpreter lock), for a bounded time."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    zero = np.zeros(n * N)

    def one(i):
        S, P, g, gl = host[i]
        O.pcg(S, P, g, zero, n, N, tol, max_iter, O.EXIT_REL)
        O.pcg(S, P, gl, zero, n, N, tol, max_iter, O.EXIT_REL)
        return 2

    done, t0 = 0, time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        while True:
            done += sum(ex.map(one, range(len(host))))
            if time.perf_counter() - t0 >= seconds:
                break
    el = time.perf_counter() - t0
    return {"value": round(done / el, 1), "unit": "fwd+adj solves/s", "cores": threads, "kind": "port",
            "host": host_cpu(),
            "sample": f"{done // 2} forward + adjoint FP64 solve pairs of the 64 n=14,N=128 systems "
                      f"(oracle/pcg_oracle.c), {threads} threads, {el:.1f} s"}


def _generate(gbdpcg, np, specs, seed0, dtype):
    """Seeded LQR Schur systems, generated on host threads (the generator releases
    the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(lambda a: gbdpcg.lqr_generate(a[1][0], a[1][0] // 2, a[1][1], seed0 + a[0], dtype=dtype),
                           enumerate(specs)))


def _graph_time(torch, body, replays):
    """Captures body(stream) in one graph, replays it once untimed, then times
    `replays` replays with events on the replay stream."""
    stream = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=stream):
        body(stream)
    graph.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    cur = torch.cuda.current_stream()
    a.record(cur)
    for _ in range(replays):
        graph.replay()
    b.record(cur)
    torch.cuda.synchronize()
    return a.elapsed_time(b) * 1000.0 / replays


def flops_per_iteration(n, N):
    """Algorithmic flops of one PCG iteration of one system: the two
    block-tridiagonal products (S p and Pinv r: 2 x 2 (3N - 2) n^2) and ten
    vector flops per entry (two dot products, three axpys)."""
    return 4 * (3 * N - 2) * n * n + 10 * n * N


def valu_roofline(flops, us, rates, dtype_name):
    """Algorithmic flops over the time, against the device's vector FMA rate: 256
    CUs x 64 lanes x 2 flops x 2.4 GHz = 78.6 TFLOP/s for FP32 without
    packed-FP32 instructions (the build disables them) and for FP64 alike."""
    tf = flops / (us * 1e-6) / 1e12
    out = {"bound": "valu", "achieved": round(tf, 3), "peak": round(VALU_F32_PEAK_TFLOPS, 2), "unit": "TFLOP/s",
           "frac": round(tf / VALU_F32_PEAK_TFLOPS, 4), "algorithmic_flops": int(flops), "dtype": dtype_name,
           "note": "matrices resident in registers (LDS for the vector windows) for the whole launch; "
                   "flops from the kernel's own per-system iteration counts"}
    if dtype_name == "f32":
        out["peak_measured"] = round(rates["valu_f32_tflops"], 2)
        out["frac_of_measured"] = round(tf / rates["valu_f32_tflops"], 4)
    return out


def cpu_systems(np, jobs, seconds, label):
    """CPU baseline over many systems: the FP64 restatement of the reference
    iteration (oracle/), one solve per task on a pool of host threads (the C
    solver releases the interpreter lock), cycling over `jobs` -- tuples
    (S, Pinv, rhs, lambda0, n, N) in FP64 -- for a bounded time."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)

    def one(j):
        S, P, rhs, l0, n, N = jobs[j]
        O.pcg(S, P, rhs, l0, n, N, 1e-5, 173, O.EXIT_REL)
        return 1

    done, t0 = 0, time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        while True:
            done += sum(ex.map(one, range(len(jobs))))
            if time.perf_counter() - t0 >= seconds:
                break
    el = time.perf_counter() - t0
    return {"value": round(done / el, 1), "unit": "solves/s", "cores": threads, "kind": "port", "host": host_cpu(),
            "sample": f"{done} FP64 solves cycling over {len(jobs)} {label} (oracle/pcg_oracle.c), {threads} threads, "
                      f"{el:.1f} s"}


def lqr_n2_N3(gbdpcg, torch, np, dev, seconds):
    """The LQR n=2, m=1, N=3 system with the block-Jacobi Pinv in FP32 (the shape
    of the reference's compile example), relative residual 1e-5, max_iter 173:
    back-to-back device solves, the host-pointer drop-in, and the host loop (the
    FP64 oracle, one thread) beside them."""
    from oracle import oracle as O
    n, m, N = 2, 1, 3
    S_h, g_h = gbdpcg.lqr_generate(n, m, N, 1000, dtype=np.float32)
    S, g = torch.from_numpy(S_h).to(dev), torch.from_numpy(g_h).to(dev)
    P = torch.empty_like(S)
    gbdpcg.precond_device(S, P, n, N, gbdpcg.PRECOND_JACOBI)
    P_h = P.cpu().numpy()
    stream = torch.cuda.current_stream()
    dres = torch.zeros(gbdpcg.RESULT_BYTES, dtype=torch.uint8, device=dev)
    out = {"workload": "LQR n=2 m=1 N=3, block-Jacobi Pinv, FP32, rel resid 1e-5, max_iter 173"}
    for name, rec in (("fused", gbdpcg.RECUR_FUSED), ("reference", gbdpcg.RECUR_REFERENCE)):
        kw = dict(tol=1e-5, max_iter=173, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=rec)
        lams = torch.zeros((220, n * N), device=dev)
        res = gbdpcg.solve(S, g, torch.zeros(n * N, device=dev), n, N, P, **kw)
        for i in range(20):
            gbdpcg.solve_async(S, g, lams[i], n, N, P, dres, **kw)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        for i in range(20, 220):
            gbdpcg.solve_async(S, g, lams[i], n, N, P, dres, **kw)
        b.record(stream)
        torch.cuda.synchronize()
        us = a.elapsed_time(b) * 1000.0 / 200
        lam = lams[219].cpu().numpy()
        exact = np.linalg.solve(O.cbtd_to_dense(S_h, n, N), g_h.astype(np.float64))
        out[name] = {"us_per_solve": round(us, 3), "iters": res.iters,
                     "rel_err_vs_fp64_dense": rel_err(np, lam, exact)}
    kw = dict(tol=1e-5, max_iter=173, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    # the host-pointer call through the C ABI, arguments prepared beforehand
    import ctypes
    cfg_h = gbdpcg.make_config(**kw)
    lh = [np.zeros(n * N, dtype=np.float32) for _ in range(110)]
    each = []
    for i in range(110):
        t0 = time.perf_counter()
        gbdpcg.lib().gbdpcg_solve_f32(S_h.ctypes.data, P_h.ctypes.data, g_h.ctypes.data, lh[i].ctypes.data, n, N,
                                      ctypes.byref(cfg_h))
        if i >= 10:
            each.append((time.perf_counter() - t0) * 1e6)
    out["host_pointer_dropin_us"] = round(float(np.mean(each)), 2)
    # the host loop: FP64 oracle, one thread, bounded sample
    S64, P64, g64, z = S_h.astype(np.float64), P_h.astype(np.float64), g_h.astype(np.float64), np.zeros(n * N)
    for _ in range(100):
        O.pcg(S64, P64, g64, z, n, N, 1e-5, 173, O.EXIT_REL)
    cnt, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(1000):
            _, it_cpu, *_ = O.pcg(S64, P64, g64, z, n, N, 1e-5, 173, O.EXIT_REL)
        cnt += 1000
    cpu_us = (time.perf_counter() - t0) * 1e6 / cnt
    out["cpu_baseline"] = {"value": round(cpu_us, 3), "unit": "us/solve", "cores": 1, "kind": "port",
                           "host": host_cpu(), "sample": f"{cnt} FP64 solves of this system "
                                                         f"(oracle/pcg_oracle.c, {it_cpu} iterations each)"}
    # latency roofline: one workgroup (N = 3 is below the register kernels' minimum of
    # four knots, so the LDS kernel runs it alone); its per-iteration time against
    # the exchange round of a single workgroup
    kw0 = dict(kw, max_iter=0)
    ev = []
    for i in range(50):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        lams[i].zero_()
        e0.record(stream)
        gbdpcg.solve_async(S, g, lams[i], n, N, P, dres, **kw)
        e1.record(stream)
        f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        f0.record(stream)
        gbdpcg.solve_async(S, g, lams[i], n, N, P, dres, **kw0)
        f1.record(stream)
        ev.append((e0, e1, f0, f1))
    torch.cuda.synchronize()
    kern = float(np.median([e0.elapsed_time(e1) * 1000.0 for e0, e1, _, _ in ev]))
    setup = float(np.median([f0.elapsed_time(f1) * 1000.0 for _, _, f0, f1 in ev]))
    per_iter = (kern - setup) / max(out["fused"]["iters"], 1)
    floor = gbdpcg.exchange_floor(1, 256, rounds=4000)["us_per_round"]
    out["roofline"] = {"bound": "latency", "kernel_us": round(kern, 3), "setup_us": round(setup, 3),
                       "per_iteration_us": round(per_iter, 3), "exchange_floor_us": round(floor, 3),
                       "frac": round(floor / per_iter, 4) if per_iter > 0 else None,
                       "note": "36 unknowns: launch and on-chip latency bound; HBM bytes per solve "
                               f"{(6 * n * n * N + 3 * n * N) * 4}"}
    out["all_converged"] = res.status == gbdpcg.CONVERGED
    out["max_rel_err_vs_fp64_cholesky"] = {"lambda": max(out["fused"]["rel_err_vs_fp64_dense"],
                                                         out["reference"]["rel_err_vs_fp64_dense"])}
    return out


def uneven_512(gbdpcg, torch, np, dev, autotune, rates, cpu_seconds):
    rng = np.random.default_rng(3000)
    specs = [(int(rng.choice([6, 12, 14, 28])), int(rng.integers(16, 257))) for _ in range(512)]
    systems = []
    gen = _generate(gbdpcg, np, specs, 3000, np.float32)
    for (n, N), (S_h, g_h) in zip(specs, gen):
        S = torch.from_numpy(S_h).to(dev)
        P = torch.empty_like(S)
        gbdpcg.precond_device(S, P, n, N, gbdpcg.PRECOND_SS)
        systems.append(dict(S=S, Pinv=P, gamma=torch.from_numpy(g_h).to(dev), lam=torch.zeros(n * N, device=dev),
                            n=n, N=N, tol=1e-5, max_iter=173))
    plan = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED, autotune=autotune)
    lams = [d["lam"] for d in systems]

    def body(st):
        torch._foreach_zero_(lams)
        plan.run(stream=st)

    us = _graph_time(torch, body, 10)
    res = plan.results()  # every replay starts from lambda0 = 0: the same counts each time
    err = max(rel_err(np, d["lam"].cpu().numpy(), banded_cholesky_solve(np, S_h, g_h, d["n"], d["N"]))
              for d, (S_h, g_h) in zip(systems, gen))
    flops = sum(r.iters * flops_per_iteration(d["n"], d["N"]) for d, r in zip(systems, res))
    # CPU baseline: the same systems' forward solves (FP64), threads over systems
    jobs = [(S_h.astype(np.float64), d["Pinv"].cpu().numpy().astype(np.float64), g_h.astype(np.float64),
             np.zeros(d["n"] * d["N"]), d["n"], d["N"]) for d, (S_h, g_h) in zip(systems, gen)]
    out = {"systems": 512, "us_per_batch": round(us, 1), "solves_per_s": round(512 / (us * 1e-6), 1),
           "mean_iters": float(np.mean([r.iters for r in res])),
           "all_converged": all(r.status == 0 for r in res),
           "max_rel_err_vs_fp64_cholesky": {"lambda": err, "systems_checked": 512},
           "roofline": valu_roofline(flops, us, rates, "f32"),
           "cpu_baseline": cpu_systems(np, jobs, cpu_seconds, "of the 512 uneven systems"),
           "plan": plan.info(), "tune": plan.tune_report}
    del jobs
    # beside it: the untuned plan, one persistent launch for every system
    plan2 = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED,
                         autotune=False, split_launches=False)

    def body2(st):
        torch._foreach_zero_(lams)
        plan2.run(stream=st)

    us2 = _graph_time(torch, body2, 10)
    res2 = plan2.results()
    out["one_launch_untuned"] = {"us_per_batch": round(us2, 1), "solves_per_s": round(512 / (us2 * 1e-6), 1),
                                 "all_converged": all(r.status == 0 for r in res2), "plan": plan2.info(),
                                 "roofline": valu_roofline(sum(r.iters * flops_per_iteration(d["n"], d["N"])
                                                               for d, r in zip(systems, res2)), us2, rates, "f32")}
    return out


def mpc_sweep(gbdpcg, torch, np, dev, dtype, rates, cpu_seconds, count=256, n=28, N=512, steps=10):
    """Ten warm-started forward + adjoint re-solves of 256 systems in one graph.
    Every timed replay starts from the same state (restored outside the timed
    region), so one eager pass of the same ten steps gives the kernel's exact
    per-step iteration counts (and must end in the same bits)."""
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    systems = []
    for i, (S_h, g_h) in enumerate(_generate(gbdpcg, np, [(n, N)] * count, 5100, dtype)):
        S = torch.from_numpy(S_h).to(dev)
        P = torch.empty_like(S)
        gbdpcg.precond_device(S, P, n, N, gbdpcg.PRECOND_SS)
        systems.append(dict(S=S, Pinv=P, gamma=torch.from_numpy(g_h).to(dev), lam=torch.zeros(n * N, dtype=tdt,
                            device=dev), n=n, N=N, tol=1e-5, max_iter=173, adjoint=True,
                            glam=torch.from_numpy(gbdpcg.randn(n * N, 9100 + i, dtype=dtype)).to(dev),
                            ggamma=torch.zeros(n * N, dtype=tdt, device=dev), gS=torch.zeros_like(S)))
        del S_h, g_h
    plan = gbdpcg.Batch(systems, exit_mode=gbdpcg.EXIT_REL_RESID, recurrence=gbdpcg.RECUR_FUSED)
    lam0 = [d["lam"].clone() for d in systems]
    gam0 = [d["gamma"].clone() for d in systems]

    def restore():
        torch._foreach_copy_([d["lam"] for d in systems], lam0)
        torch._foreach_copy_([d["gamma"] for d in systems], gam0)
        torch.cuda.synchronize()

    stream = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=stream):
        for k in range(steps):
            plan.warm_start(seed=900 + k, noise_scale=1e-2, stream=stream)
            plan.run(stream=stream)
    times = []
    cur = torch.cuda.current_stream()  # graph.replay() launches on the current stream
    for rep in range(4):  # the first replay untimed
        restore()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(cur)
        graph.replay()
        b.record(cur)
        torch.cuda.synchronize()
        if rep:
            times.append(a.elapsed_time(b) * 1000.0)
    us = float(np.median(times))
    replay_final = [(d["lam"].cpu().numpy().copy(), d["ggamma"].cpu().numpy().copy()) for d in systems[::37]]
    # the same ten steps eagerly: per-step iteration counts, and the last step's inputs
    restore()
    iters_total, last_in = 0, None
    for k in range(steps):
        plan.warm_start(seed=900 + k, noise_scale=1e-2)
        if k == steps - 1:
            last_in = [d["lam"].cpu().numpy().astype(np.float64) for d in systems[:16]]
        plan.run()
        res = plan.results()
        iters_total += sum(r.iters + r.adj_iters for r in res)
    same = all(np.array_equal(d["lam"].cpu().numpy(), l) and np.array_equal(d["ggamma"].cpu().numpy(), gg)
               for d, (l, gg) in zip(systems[::37], replay_final))
    e_lam = e_gg = 0.0
    for d in systems[::count // 16]:  # 16 systems, final state of the last re-solve
        S64 = d["S"].cpu().numpy()
        x = banded_cholesky_solve(np, S64, d["gamma"].cpu().numpy(), n, N)
        y = banded_cholesky_solve(np, S64, d["glam"].cpu().numpy(), n, N)
        e_lam = max(e_lam, rel_err(np, d["lam"].cpu().numpy(), x))
        e_gg = max(e_gg, rel_err(np, d["ggamma"].cpu().numpy(), y))
    # CPU baseline: the last re-solve of 16 systems (forward from its warm start,
    # adjoint from zero), FP64, threads over solves
    jobs = []
    for d, l_in in zip(systems[:16], last_in):
        S64 = d["S"].cpu().numpy().astype(np.float64)
        P64 = d["Pinv"].cpu().numpy().astype(np.float64)
        jobs.append((S64, P64, d["gamma"].cpu().numpy().astype(np.float64), l_in, n, N))
        jobs.append((S64, P64, d["glam"].cpu().numpy().astype(np.float64), np.zeros(n * N), n, N))
    cpu = cpu_systems(np, jobs, cpu_seconds, "forward and adjoint solves of the last re-solve of 16 systems")
    dt_name = "f64" if dtype == np.float64 else "f32"
    out = {"systems": count, "re_solves": steps, "ms_per_graph": round(us / 1000.0, 2),
           "fwd_adj_solves_per_s": round(count * steps / (us * 1e-6), 1),
           "mean_iters_per_solve": round(iters_total / (2 * count * steps), 2),
           "last_mean_iters_fwd": float(np.mean([r.iters for r in res])),
           "last_mean_iters_adj": float(np.mean([r.adj_iters for r in res])),
           "all_converged": all(r.status == 0 and r.adj_status == 0 for r in res),
           "replay_equals_eager": bool(same),
           "max_rel_err_vs_fp64_cholesky": {"lambda": e_lam, "g_gamma": e_gg,
                                             "systems_checked": len(systems[::count // 16])},
           "roofline": valu_roofline(iters_total * flops_per_iteration(n, N), us, rates, dt_name),
           "cpu_baseline": dict(cpu, value=round(cpu["value"] / 2, 1), unit="fwd+adj solves/s"),
           "plan": plan.info()}
    plan.close()
    del systems, lam0, gam0
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
