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
the dominant kernel (algorithmic bytes / its HIP-event duration, the event
interval less an empty event pair's) and the CPU oracle's throughput on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

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


# HBM bytes per launch of the dominant kernel, measured with rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE (scaled by a plain device copy of known size), by
# (workload, precision, K, form).  Read from the profile named in
# "traffic_source"; a form without a measured profile reports traffic null.
PMC_TRAFFIC = {
    ("worker_firm_1e7", "fp64", 1, "csr"): ({"k_fused": 324.1e6, "k_dtx_csr": 311.1e6, "k_update": 88.1e6},
                                            "profiles/worker_firm_1e7_fp64_k1_csr.md"),
    ("worker_firm_1e7", "fp64", 1, "pieces"): ({"k_fused": 332.4e6, "k_dtx_csr": 103.3e6, "k_update": 88.1e6},
                                               "profiles/worker_firm_1e7_fp64_k1.md"),
    ("worker_firm_1e7_k32", "fp64", 32, "pieces"): ({"k_fused": 6065.9e6, "k_dtx_csr": 1379.6e6,
                                                     "k_update": 2554.9e6}, "profiles/worker_firm_1e7_fp64_k32.md"),
}


def kernel_bytes(N, n, kinds, K, s, factored, weighted, plan):
    """Algorithmic HBM bytes per launch of each kernel of one LSMR iteration
    (each array element counted once; gathers counted as their distinct
    elements), for the plan the library built (fehip_plan_info: few-level FEs
    in the fused pass or not, the piece form or the CSR pass).  factored: no
    FE is interacted, so the fused pass reads one sqrt(w) stream (none
    without weights) instead of every FE's cache."""
    G = len(n)
    p = sum(n)
    gs = kinds.index("sorted")
    fused_dense = bool(plan["fused_dense"])
    npc = plan["pieces"]
    ndense = kinds.count("dense")
    out = {}
    # fused pass: refs of every FE but s (whose level comes from head bits and
    # their word prefixes; with G = 3 and the few-level FE in the pass, the
    # other two share one packed stream), caches (or sqrt(w)), u read + write,
    # each gathered coefficient once, the FE s (or piece) level map, the raw
    # sums out (FE s levels in doubles, or the pieces in T)
    nref = G - 2 if (G == 3 and ndense == 1 and fused_dense) else G - 1
    per_obs = N * nref * 4 + ((N * s if weighted else 0) if factored else N * G * s)
    if npc:
        out["k_fused"] = per_obs + K * N * 2 * s + K * p * s + N // 8 + N // 16 + 4 * npc + s * K * npc
    else:
        out["k_fused"] = per_obs + K * N * 2 * s + K * p * s + N // 8 + N // 16 + 4 * n[gs] + 8 * K * n[gs]
    csr = 0
    dense = 0
    scaled_u = factored and K < 8  # the CSR pass then reads the stored sqrt(w) .* u, no cache
    for g in range(G):
        if kinds[g] == "csr" and not npc:
            csr += 4 * N + (0 if scaled_u else N * s) + K * N * s + 2 * K * n[g] * s + N // 8 + 4 * n[g]
        elif kinds[g] == "dense" and not fused_dense:
            dense += N * (4 + s) + K * N * s
    if npc:
        # the piece launch: for FE s (identity order) and FE t (through the
        # pieces' permutation by FE t level): the order and each position's
        # level (4 + 4 bytes a piece), the piece sums, head bits, the level's
        # scale, and each level's v read and written
        gt = [g for g in range(G) if kinds[g] == "csr"][0]
        for g in (gs, gt):
            csr += 8 * npc + K * npc * s + npc // 8 + n[g] * s + 2 * K * n[g] * s
        out["k_dtx_csr"] = csr
        out["k_dtx_fixup"] = 0
    else:
        # FE s levels: raw sum (double) + v read + write, one flag byte (+ scale).
        # With CSR FEs these items run as the first blocks of the CSR launch
        fix = K * n[gs] * (8 + 2 * s) + n[gs] + (n[gs] * s if factored else 0)
        if csr:
            out["k_dtx_csr"] = csr + fix
            out["k_dtx_fixup"] = 0
        else:
            out["k_dtx_csr"] = 0
            out["k_dtx_fixup"] = fix
    out["k_dtx_dense"] = dense
    # v, h, hbar, x read + write (+ scale read, scale .* v written)
    out["k_update"] = 8 * K * p * s + ((K + 1) * p * s if factored else 0)
    # the pair formula of the survey: D*v + D'u over all G FEs with perm+cache each
    out["pair"] = 2 * N * G * (4 + s) + K * (2 * N * s + N * G * s + 3 * p * s) + 4 * (p + G)
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(panel, column, tol, maxiter, nthreads):
    """The oracle (C restatement of the reference's CPU path) solving one
    column of the workload to convergence: set-up, then solve_residuals! at
    the bench's tol.  Returns the baseline entry and the oracle's residual."""
    import oracle

    y = np.ascontiguousarray(panel.Y[:, column])
    t0 = time.perf_counter()
    s = oracle.OracleSolver(panel.refs, panel.n, panel.interactions, panel.weights, True, nthreads)
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    r, iters, conv = s.solve_residuals(y.copy(), tol=tol, maxiter=maxiter)
    solve = time.perf_counter() - t0
    entry = {
        "value": 1.0 / solve,
        "unit": "columns/s",
        "cores": nthreads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": (f"oracle (C restatement of src/SolverCPU.jl + src/utils/lsmr.jl + "
                   f"src/AbstractFixedEffectSolver.jl; D'x serial as the reference, D*x and cache threaded in "
                   f"100k-row chunks) solving column {column} of the same panel to tol {tol:g}: set-up "
                   f"{setup:.2f} s (not in value, as for the GPU), solve {solve:.2f} s, {iters - 1} LSMR iterations"),
        "setup_s": setup,
        "solve_s": solve,
        "lsmr_iters": int(iters) - 1,
        "converged": bool(conv),
    }
    return entry, r


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(float(np.linalg.norm(b)), 1e-300))


# the last slot is an empty event pair: the timing's own overhead, which every
# other slot includes
KERNEL_NAMES = ["k_fused", "k_fin_db", "k_dtx_csr", "k_dtx_dense", "k_dtx_fixup", "k_fin_dtx", "k_update",
                "event_pair"]


def kernel_report(feM, panel, K, precision, name, form_key=None):
    """Per-kernel HIP-event times of one iteration (fehip_time_kernels on the
    handle's stream), the dominant kernel's roofline entry, and the survey's
    pair-formula bytes over the iteration's D*x + D'u launches."""
    import ctypes as C

    from fixedeffects_rocm._lib import check, lib

    ms = np.zeros(8)
    check(lib.fehip_time_kernels(feM._h, K, 20, ms.ctypes.data_as(C.POINTER(C.c_double))))
    s_max = int(np.argmax(panel.n))
    kinds = ["sorted" if g == s_max else ("dense" if k <= 64 else "csr") for g, k in enumerate(panel.n)]
    factored = all(z is None for z in panel.interactions)
    # the FP32 streams of the mixed mode's first phase are what the timing hook runs
    fp32 = precision != "fp64"
    plan = feM.plan_info(K, fp32_storage=fp32)
    s = 4 if fp32 else 8
    kb = kernel_bytes(panel.N, list(panel.n), kinds, K, s, factored, panel.weights is not None, plan)
    kt = {nm: float(m) for nm, m in zip(KERNEL_NAMES, ms)}
    dom = max(["k_fused", "k_dtx_csr", "k_dtx_dense", "k_dtx_fixup", "k_update"], key=lambda q: kt[q])
    # the launch's duration: its event interval less the interval of an empty
    # event pair recorded the same way on the same stream (the timing's own
    # cost, which every interval includes)
    ovh = kt["event_pair"] if 0.0 < kt["event_pair"] < 0.5 * kt[dom] else 0.0
    dur_ms = kt[dom] - ovh
    achieved = kb[dom] / (dur_ms * 1e-3) / 1e9
    pair_ms = sum(kt[q] for q in KERNEL_NAMES[:6])
    form = "pieces" if plan["pieces"] else "csr"
    traffic, src = PMC_TRAFFIC.get((name, precision, K, form), ({}, None))
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic.get(dom), "traffic_source": src if traffic else None,
            "algorithmic_bytes": int(kb[dom]), "duration_us": dur_ms * 1e3, "event_interval_us": kt[dom] * 1e3,
            "event_pair_us": ovh * 1e3}
    return {"kernel_ms": kt, "roofline": roof, "hbm_gbs_pair": kb["pair"] / (pair_ms * 1e-3) / 1e9,
            "plan": plan, "form": form}


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


def measure(feM, pristine, work, K, N, steps, warmup, tol, maxiter, label):
    """Warm-up, then `steps` timed solves of K device-resident columns (each
    step restores the columns first, inside the timed region, as a copy in
    HBM)."""
    from fixedeffects_rocm import device

    def step():
        work.copy_from(pristine)
        device.synchronize()
        return feM.solve_residuals_device(work.address, K, N, tol=tol, maxiter=maxiter)

    for w in range(warmup):
        its, cvs = step()
        log(f"[bench] {label} warmup {w}: iters {its[:4]}{'...' if K > 4 else ''} converged {all(cvs)}")
    device.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        its, cvs = step()
    device.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    log(f"[bench] {label}: {steps} steps, {ms:.2f} ms/step, iters {its[:4]}")
    return ms, [int(i) - 1 for i in its], [bool(c) for c in cvs]


def batched_leg(args, precision, plain_pair_gbs, plain_value):
    """The 32-column workload (the 1e7 panel with 31 regressors) in one
    batched solve, in FP64 or the mixed mode, beside the K = 1 value."""
    from fixedeffects_rocm import api, device, synth

    name = "worker_firm_1e7_k32"
    panel = _panel_cache.get(name)
    if panel is None:
        t0 = time.perf_counter()
        panel = synth.config(name)
        _panel_cache[name] = panel
        log(f"[bench] panel {name}: N={panel.N} K={panel.K} ({time.perf_counter() - t0:.1f} s)")
    K, N = panel.K, panel.N
    fes = [api.FixedEffect.from_refs(r, k, z) for r, k, z in zip(panel.refs, panel.n, panel.interactions)]
    feM = api.FixedEffectSolverROCm(fes, panel.weights, precision)
    sched = None
    if not args.no_autotune:
        sched = feM.autotune(K=K, reps=3)
    pristine = device.DeviceBuffer(N * K)
    pristine.upload(np.asfortranarray(panel.Y).ravel(order="F"))
    work = device.DeviceBuffer(N * K)
    ms, its, cvs = measure(feM, pristine, work, K, N, args.leg_steps, 1, args.tol, args.maxiter,
                           f"K=32 {precision}")
    r = work.download(N * K)
    rep = kernel_report(feM, panel, K, precision, name)
    feM.close()
    del pristine, work
    leg = {"workload": name, "precision": precision, "K": K, "value": K / (ms / 1e3), "ms_per_step": ms,
           "vs_value": K / (ms / 1e3) / plain_value,
           "lsmr_iters": {"min": min(its), "max": max(its), "mean": float(np.mean(its))},
           "converged": all(cvs), "schedule": sched, "kernel_ms": rep["kernel_ms"], "roofline": rep["roofline"],
           "hbm_gbs_pair": rep["hbm_gbs_pair"], "hbm_gbs_vs_plain": rep["hbm_gbs_pair"] / plain_pair_gbs,
           "form": rep["form"]}
    return leg, r, panel


_panel_cache = {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="worker_firm_1e7", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle's solves (and the parity fields)")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--maxiter", type=int, default=10000)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--pcie-steps", type=int, default=2, help="timed solves through the host-buffer entry point")
    ap.add_argument("--no-autotune", action="store_true", help="keep the occupancy-sized default schedule")
    ap.add_argument("--panels", type=int, default=256, help="panels of the uneven_panels workload")
    ap.add_argument("--alt-steps", type=int, default=3, help="timed steps of the projected LSMR beside `value`")
    ap.add_argument("--leg-steps", type=int, default=2, help="timed steps of each 32-column leg (0: no legs)")
    ap.add_argument("--algorithm", default="lsmr", choices=["lsmr", "projected"],
                    help="lsmr: the reference's LSMR; projected: LSMR with the FE of most levels eliminated")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; the solve is not sharded across GPUs "
                                   "(bench.py --gpus 1)"}))
        sys.exit(2)

    from fixedeffects_rocm import api, device, synth

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
    ms_per_step, its, cvs = measure(feM, pristine, work, K, N, args.steps, args.warmup, args.tol, args.maxiter,
                                    "value")
    value = K / (ms_per_step / 1e3)
    r = work.download(N)

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
    if args.pcie_steps:
        log(f"[bench] host-buffer path: {pcie_ms:.2f} ms/step")

    # normal equations of the result: ||A'(sqrt(w) r)||_inf / ||sqrt(w) y||_2 (size-independent check)
    sw = np.sqrt(panel.weights) if panel.weights is not None else 1.0
    g = feM.mul_At(sw * r)
    neq = float(np.abs(g).max() / np.linalg.norm(sw * panel.Y[:, 0]))

    # the same workload with the projected LSMR (the FE of most levels
    # eliminated): same least-squares residual, fewer iterations; reported
    # beside `value`, which stays the reference algorithm's
    alt, rp = None, None
    if args.algorithm == "lsmr" and args.alt_steps > 0:
        feM.set_algorithm("projected")
        alt_ms, its_p, cvs_p = measure(feM, pristine, work, K, N, args.alt_steps, 1, args.tol, args.maxiter,
                                       "projected")
        rp = work.download(N)
        feM.set_algorithm("lsmr")
        alt = {"algorithm": "projected", "value": K / (alt_ms / 1e3), "ms_per_step": alt_ms,
               "lsmr_iters": its_p, "converged": all(cvs_p), "rel_diff_vs_value_run": rel(rp, r)}

    rep = kernel_report(feM, panel, K, args.precision, name)
    feM.close()
    del pristine, work
    iters = int(np.max(its))

    # CPU oracle: a converged solve of column 0, and the parity of the GPU's
    # residuals against it (same tol)
    cpu, parity = None, None
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if not args.no_cpu:
        cpu, ro = cpu_baseline(panel, 0, args.tol, args.maxiter, nthreads)
        log(f"[bench] cpu baseline {cpu['value']:.4g} columns/s ({cpu['solve_s']:.2f} s, {cpu['lsmr_iters']} iters)")
        parity = {"rel_err_vs_oracle": rel(r, ro), "iters_gpu": int(its[0]), "iters_oracle": cpu["lsmr_iters"],
                  "note": "column 0 of `value`'s last step against the oracle's solve at the same tol"}
        if rp is not None:
            parity["projected_rel_err_vs_oracle"] = rel(rp, ro)

    legs = {}
    if args.leg_steps > 0 and name == "worker_firm_1e7" and args.precision == "fp64" and args.algorithm == "lsmr":
        leg64, r64, p32 = batched_leg(args, "fp64", rep["hbm_gbs_pair"], value)
        legmx, rmx, _ = batched_leg(args, "mixed", rep["hbm_gbs_pair"], value)
        Nb = p32.N
        legmx["rel_diff_vs_fp64_leg"] = [rel(rmx[c * Nb:(c + 1) * Nb], r64[c * Nb:(c + 1) * Nb])
                                         for c in range(p32.K)]
        if not args.no_cpu:
            o32, ro32 = cpu_baseline(p32, 0, args.tol, args.maxiter, nthreads)
            leg64["col0_rel_err_vs_oracle"] = rel(r64[:Nb], ro32)
            legmx["col0_rel_err_vs_oracle"] = rel(rmx[:Nb], ro32)
            leg64["col0_iters_oracle"] = legmx["col0_iters_oracle"] = o32["lsmr_iters"]
        legs = {"batched_k32": leg64, "mixed_k32": legmx}

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
        "lsmr_iters": its,
        "converged": all(cvs),
        "normal_eq_note": "max_level |A'(sqrt(w) r)| / ||sqrt(w) y|| of column 0",
        "normal_eq_rel": neq,
        "parity": parity,
        "setup_s": setup_s,
        "schedule": sched,
        "form": rep["form"],
        "kernel_ms": rep["kernel_ms"],
        "hbm_gbs_pair": rep["hbm_gbs_pair"],
        "roofline": rep["roofline"],
        "cpu_baseline": cpu,
        "projected_lsmr": alt,
        "pcie_inclusive": {"value": K / (pcie_ms / 1e3) if args.pcie_steps else None, "ms_per_step": pcie_ms,
                           "note": "fehip_solve_residuals on host columns: H2D + solve + D2H (not `value`)"},
    }
    out.update(legs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
