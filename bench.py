#!/usr/bin/env python3
"""Throughput of the fixed-effect residualization on one MI355X.

Workload (default): worker / firm / year panel, N = 10,000,000 observations,
3 fixed effects with 1,000,000 Zipf(1.1)-sized worker levels, 100,000 firm
levels (mobility 0.05) and 20 year levels, LogNormal(0, 0.5) weights, one
column, Float64, tol = 1e-8, maxiter = 10000 -- the plain FP64 GPU drop-in.

A step = one residualization of the column(s) through the C-ABI entry point
that takes device-resident columns (fehip_solve_residuals_device): weighting
and reordering in, the whole LSMR loop, the final r = b - A x and the
reordering out.  The solver handle (refs, CSR bucketings, scales, caches) is
built once before timing, as FixedEffectModels reuses one solver for all
columns.  Inputs are in HBM when the timed region starts.

Printed: one JSON line with columns/s, LSMR iterations, a roofline entry for
the dominant kernel (algorithmic bytes / HIP-event duration) and the CPU
oracle's throughput on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fixedeffects.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "worker_firm_1e7": ("worker_firm_1e7", 1),
    "worker_firm_1e7_k32": ("worker_firm_1e7_k32", 32),
    "worker_firm_1e8_k16": ("worker_firm_1e8_k16", 16),
    "uneven_panels": ("uneven_panels", None),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (scaled by a plain device copy of known size), by (workload, precision):
# profiles/worker_firm_1e7_fp64_k1.md
PMC_TRAFFIC = {
    ("worker_firm_1e7", "fp64"): {"k_fused": 324.1e6, "k_dtx_csr": 311.1e6, "k_update": 88.1e6},
}


def kernel_bytes(N, G, n, kinds, K, s, factored, weighted):
    """Algorithmic HBM bytes per launch of each kernel of one LSMR iteration
    (each array element counted once; gathers counted as their distinct
    elements).  factored: no FE is interacted, so the fused pass reads one
    sqrt(w) stream (none without weights) instead of every FE's cache."""
    p = sum(n)
    gs = kinds.index("sorted")
    ndense_lev = sum(k for k, kd in zip(n, kinds) if kd == "dense")
    fused_dense = 0 < ndense_lev <= 32
    out = {}
    # fused pass: refs of every FE but s (whose level comes from its head bits
    # and their word prefixes; with G = 3 and one few-level FE the other two
    # share one packed stream), caches (or sqrt(w)), u read + write, each
    # coefficient once, raw FE s sums (double) out
    nref = G - 2 if (G == 3 and kinds.count("dense") == 1 and fused_dense) else G - 1
    per_obs = N * G * s + N * nref * 4 if not factored else N * nref * 4 + (N * s if weighted else 0)
    out["k_fused"] = per_obs + K * N * 2 * s + K * p * s + N // 8 + N // 16 + 8 * K * n[gs]
    csr = 0
    dense = 0
    for g in range(G):
        if kinds[g] == "csr":
            csr += 4 * N + N * s + K * N * s + 2 * K * n[g] * s + N // 8 + 4 * n[g]
        elif kinds[g] == "dense" and not fused_dense:
            dense += N * (4 + s) + K * N * s
    out["k_dtx_csr"] = csr
    out["k_dtx_dense"] = dense
    # FE s levels: raw sum (double) + v read + write, one flag byte (+ scale).
    # With CSR FEs these items run as the first blocks of the CSR launch
    # (k_dtx_seg_fix), whose slot then carries their bytes
    fix = K * n[gs] * (8 + 2 * s) + n[gs] + (n[gs] * s if factored else 0)
    if csr:
        out["k_dtx_csr"] += fix
        out["k_dtx_fixup"] = 0
    else:
        out["k_dtx_fixup"] = fix
    # v, h, hbar, x read + write (+ scale read, scale .* v written)
    out["k_update"] = 8 * K * p * s + ((K + 1) * p * s if factored else 0)
    # the pair formula of the survey: D*v + D'u over all G FEs with perm+cache each
    out["pair"] = 2 * N * G * (4 + s) + K * (2 * N * s + N * G * s + 3 * p * s) + 4 * (p + G)
    return out


def cpu_baseline(panel, iters_gpu, budget_s, nthreads):
    """The oracle (C restatement of the reference's CPU path) timed on a
    bounded sample of the same workload."""
    import numpy as np

    import oracle

    y = np.ascontiguousarray(panel.Y[:, 0])
    t0 = time.perf_counter()
    s = oracle.OracleSolver(panel.refs, panel.n, panel.interactions, panel.weights, True, nthreads)
    setup = time.perf_counter() - t0
    b = y * (np.sqrt(panel.weights) if panel.weights is not None else 1.0)
    x0 = s.mul_At(b)
    t0 = time.perf_counter()
    s.lsmr(x0, b, atol=0.0, btol=0.0, conlim=0.0, maxiter=2)
    per_iter = (time.perf_counter() - t0) / 2
    k = int(max(2, min(200, (budget_s - setup) / max(per_iter, 1e-6))))
    t0 = time.perf_counter()
    s.lsmr(x0, b, atol=0.0, btol=0.0, conlim=0.0, maxiter=k)
    per_iter = (time.perf_counter() - t0) / k
    col_s = (iters_gpu + 1) * per_iter
    return {
        "value": 1.0 / col_s,
        "unit": "columns/s",
        "cores": nthreads,
        "kind": "port",
        "sample": (f"oracle (C restatement of src/SolverCPU.jl + src/utils/lsmr.jl; D'x serial as the reference, "
                   f"D*x and cache threaded in 100k-row chunks) on the same panel: set-up {setup:.2f} s, "
                   f"{k} LSMR iterations timed at {per_iter * 1e3:.1f} ms each; columns/s = 1 / "
                   f"((iters + 1) x per-iteration time) with iters = {iters_gpu} from the GPU run"),
        "setup_s": setup,
        "ms_per_lsmr_iter": per_iter * 1e3,
    }


def run_panels(args, precision):
    """Many uneven panels (N log-uniform in [1e4, 2e6], 2-3 FEs, 1-8 columns
    each) residualized in one fehip_solve_residuals_panels call; the same
    panels solved one after another for comparison."""
    import numpy as np

    from fixedeffects_rocm import api, device, synth

    t0 = time.perf_counter()
    panels = [synth.uneven_panel(500_000 + i, 1e4, 2e6) for i in range(args.panels)]
    log(f"[bench] {len(panels)} panels, {sum(p.N for p in panels)} obs, {sum(p.K for p in panels)} columns "
        f"({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    solvers = [api.FixedEffectSolverROCm([api.FixedEffect.from_refs(r, k, z) for r, k, z in
                                          zip(p.refs, p.n, p.interactions)], p.weights, precision,
                                         algorithm=args.algorithm) for p in panels]
    setup_s = time.perf_counter() - t0
    pristine, work = [], []
    for p in panels:
        b = device.DeviceBuffer(p.N * p.K)
        b.upload(np.asfortranarray(p.Y).ravel(order="F"))
        pristine.append(b)
        work.append(device.DeviceBuffer(p.N * p.K))
    cols = sum(p.K for p in panels)
    Ks, lds = [p.K for p in panels], [p.N for p in panels]

    def reset():
        for a, b in zip(pristine, work):
            b.copy_from(a)
        device.synchronize()

    def together():
        return api.solve_residuals_panels(solvers, [b.address for b in work], Ks, lds, tol=args.tol,
                                          maxiter=args.maxiter)

    def one_by_one():
        for s, b, p in zip(solvers, work, panels):
            s.solve_residuals_device(b.address, p.K, p.N, tol=args.tol, maxiter=args.maxiter)

    for _ in range(args.warmup):
        reset()
        together()
        reset()
        one_by_one()
    tt = ts = 0.0
    for _ in range(args.steps):
        reset()
        t0 = time.perf_counter()
        its, cvs = together()
        tt += time.perf_counter() - t0
        reset()
        t0 = time.perf_counter()
        one_by_one()
        ts += time.perf_counter() - t0
    ms = tt / args.steps * 1e3
    ms_seq = ts / args.steps * 1e3
    log(f"[bench] panels together {ms:.1f} ms/step, one by one {ms_seq:.1f} ms/step")
    flat = [i - 1 for a in its for i in a]
    out = {
        "metric": "residualized columns/s @1e-8, uneven panels (N log-uniform 1e4..2e6, 2-3 FEs, 1-8 columns)",
        "value": cols / (ms / 1e3), "unit": "columns/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded numpy PCG64 uneven worker/firm panels)",
        "config": {"workload": "uneven_panels", "panels": len(panels), "N_total": int(sum(p.N for p in panels)),
                   "columns": cols, "tol": args.tol, "precision": precision},
        "one_by_one_value": cols / (ms_seq / 1e3), "lsmr_iters_mean": float(np.mean(flat)),
        "converged": bool(all(all(c) for c in cvs)), "setup_s": setup_s,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))
    for s in solvers:
        s.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="worker_firm_1e7", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--maxiter", type=int, default=10000)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--pcie-steps", type=int, default=2, help="timed solves through the host-buffer entry point")
    ap.add_argument("--no-autotune", action="store_true", help="keep the occupancy-sized default schedule")
    ap.add_argument("--panels", type=int, default=64, help="panels of the uneven_panels workload")
    ap.add_argument("--alt-steps", type=int, default=3, help="timed steps of the projected LSMR beside `value`")
    ap.add_argument("--algorithm", default="lsmr", choices=["lsmr", "projected"],
                    help="lsmr: the reference's LSMR; projected: LSMR with the FE of most levels eliminated")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; the solve is not sharded across GPUs "
                                   "(bench.py --gpus 1)"}))
        sys.exit(2)

    import ctypes as C

    import numpy as np

    from fixedeffects_rocm import api, device, synth
    from fixedeffects_rocm._lib import check, lib

    name, K = WORKLOADS[args.workload]
    if name == "uneven_panels":
        run_panels(args, args.precision)
        return
    t0 = time.perf_counter()
    panel = synth.config(name)
    log(f"[bench] panel {name}: N={panel.N} n={panel.n} K={panel.K} ({time.perf_counter() - t0:.1f} s)")
    fes = [api.FixedEffect.from_refs(r, k, z) for r, k, z in zip(panel.refs, panel.n, panel.interactions)]
    t0 = time.perf_counter()
    feM = api.FixedEffectSolverROCm(fes, panel.weights, args.precision, algorithm=args.algorithm)
    setup_s = time.perf_counter() - t0
    log(f"[bench] solver set-up {setup_s:.2f} s")
    sched = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        sched = feM.autotune(K=K, reps=5)
        sched["seconds"] = time.perf_counter() - t0
        log(f"[bench] autotune: {sched}")

    N = panel.N
    pristine = device.DeviceBuffer(N * K)   # column c at offset c*N
    pristine.upload(np.asfortranarray(panel.Y).ravel(order="F"))
    work = device.DeviceBuffer(N * K)

    def step():
        work.copy_from(pristine)
        device.synchronize()
        return feM.solve_residuals_device(work.address, K, N, tol=args.tol, maxiter=args.maxiter)

    for w in range(args.warmup):
        its, cvs = step()
        log(f"[bench] warmup {w}: iters {its[:4]}{'...' if K > 4 else ''} converged {all(cvs)}")
    device.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        its, cvs = step()
    device.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = K / (ms_per_step / 1e3)
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step, iters {its[:4]}")

    # the same solve through the host-buffer entry point (PCIe transfers of
    # the columns included); reported beside `value`, never as it
    host_cols = [np.ascontiguousarray(panel.Y[:, c]) for c in range(K)]
    pcie_s = 0.0
    for _ in range(args.pcie_steps):
        cols = [c.copy() for c in host_cols]
        t1 = time.perf_counter()
        feM.solve_residuals_many(cols, tol=args.tol, maxiter=args.maxiter)
        pcie_s += time.perf_counter() - t1
    pcie_ms = pcie_s / max(args.pcie_steps, 1) * 1e3
    log(f"[bench] host-buffer path: {pcie_ms:.2f} ms/step")

    # normal equations of the result: ||A'(sqrt(w) r)||_inf / ||sqrt(w) y||_2 (size-independent check)
    r = work.download(N)
    sw = np.sqrt(panel.weights) if panel.weights is not None else 1.0
    g = feM.mul_At(sw * r)
    neq = float(np.abs(g).max() / np.linalg.norm(sw * panel.Y[:, 0]))

    # the same workload with the projected LSMR (the FE of most levels
    # eliminated): same least-squares residual, fewer iterations; reported
    # beside `value`, which stays the reference algorithm's
    alt = None
    if args.algorithm == "lsmr" and args.alt_steps > 0:
        feM.set_algorithm("projected")
        step()
        device.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.alt_steps):
            its_p, cvs_p = step()
        device.synchronize()
        alt_ms = (time.perf_counter() - t1) / args.alt_steps * 1e3
        rp = work.download(N)
        feM.set_algorithm("lsmr")
        alt = {"algorithm": "projected", "value": K / (alt_ms / 1e3), "ms_per_step": alt_ms,
               "lsmr_iters": [int(i) - 1 for i in its_p], "converged": bool(all(cvs_p)),
               "rel_diff_vs_value_run": float(np.linalg.norm(rp - r) / np.linalg.norm(r))}
        log(f"[bench] projected: {alt_ms:.2f} ms/step, iters {its_p[:4]}")

    # per-kernel durations (HIP events on the solver's stream)
    ms = np.zeros(8)
    check(lib.fehip_time_kernels(feM._h, K, 20, ms.ctypes.data_as(C.POINTER(C.c_double))))
    names = ["k_fused", "k_fin_db", "k_dtx_csr", "k_dtx_dense", "k_dtx_fixup", "k_fin_dtx", "k_update", "k_fin_update"]
    kinds = []
    s_max = int(np.argmax(panel.n))
    for gi, k in enumerate(panel.n):
        kinds.append("sorted" if gi == s_max else ("dense" if k <= 64 else "csr"))
    factored = all(z is None for z in panel.interactions)
    kb = kernel_bytes(N, len(panel.n), panel.n, kinds, K, 8 if args.precision == "fp64" else 4, factored,
                      panel.weights is not None)
    kt = {nm: float(m) for nm, m in zip(names, ms)}
    dom = max(["k_fused", "k_dtx_csr", "k_dtx_dense", "k_dtx_fixup", "k_update"], key=lambda q: kt[q])
    achieved = kb[dom] / (kt[dom] * 1e-3) / 1e9
    pair_ms = sum(kt[q] for q in names[:6])
    pair_gbs = kb["pair"] / (pair_ms * 1e-3) / 1e9
    iters = int(np.max(its)) - 1  # LSMR iterations (solve_residuals! reports lsmr + 1)

    cpu = None
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = cpu_baseline(panel, iters, args.cpu_seconds, nthreads)
        log(f"[bench] cpu baseline {cpu['value']:.4g} columns/s ({cpu['ms_per_lsmr_iter']:.1f} ms/iter)")

    out = {
        "metric": "residualized columns/s @1e-8, N=1e7 3 FEs; LSMR iters; HBM GB/s vs plain",
        "value": value,
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded numpy PCG64 worker/firm/year panel)",
        "config": {"workload": name, "N": N, "levels": [int(k) for k in panel.n], "K": K, "tol": args.tol,
                   "weights": "LogNormal(0,0.5)", "precision": args.precision,
                   "mode": ("plain LSMR, reference start" if args.precision == "fp64"
                            else "FP32-stream LSMR phase, then FP64 LSMR iterations from its solution")
                   + ("" if args.algorithm == "lsmr" else "; projected LSMR (FE of most levels eliminated)"),
                   "algorithm": args.algorithm},
        "lsmr_iters": [int(i) - 1 for i in its],
        "normal_eq_note": "max_level |A'(sqrt(w) r)| / ||sqrt(w) y|| of column 0",
        "converged": bool(all(cvs)),
        "normal_eq_rel": neq,
        "setup_s": setup_s,
        "schedule": sched,
        "kernel_ms": kt,
        "hbm_gbs_pair": pair_gbs,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": PMC_TRAFFIC.get((name, args.precision), {}).get(dom),
                     "algorithmic_bytes": int(kb[dom])},
        "cpu_baseline": cpu,
        "projected_lsmr": alt,
        "pcie_inclusive": {"value": K / (pcie_ms / 1e3) if args.pcie_steps else None, "ms_per_step": pcie_ms,
                           "note": "fehip_solve_residuals on host columns: H2D + solve + D2H (not `value`)"},
    }
    print(json.dumps(out))
    feM.close()


if __name__ == "__main__":
    main()
