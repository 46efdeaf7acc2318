"""Train the warm-start network on solver-generated data and evaluate it.

    python -m fixedeffects_rocm.train_warmstart --panels 2000 --eval 64

1. Training set: `--panels` seeded uneven panels (N log-uniform in
   [--n-min, --n-max]); per panel one column; features and targets from
   warmstart.training_rows (the solver's own LSMR solutions).
2. Training: fehip_mlp_fit on the GPU (bf16 MFMA forward/backward, Adam).
3. Evaluation on held-out panels (disjoint seeds, N log-uniform in
   [1e4, --eval-n-max]), every column: LSMR iterations to tol from the
   reference start (x0 = A'b), from the alternating-projection start, and from
   the learned start; residual agreement with the reference-start solve.
4. Writes the weights (npz) and a JSON report.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import api, synth, warmstart


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def evaluate(net, seeds, n_max, tol=1e-8, log_every=8):
    rows = []
    for i, sd in enumerate(seeds):
        p = synth.uneven_panel(sd, 1e4, n_max)
        fes = [api.FixedEffect.from_refs(r, k, z) for r, k, z in zip(p.refs, p.n, p.interactions)]
        with api.FixedEffectSolverROCm(fes, p.weights, "fp64") as s:
            K = p.K
            ref = [np.ascontiguousarray(p.Y[:, c]).copy() for c in range(K)]
            t0 = time.perf_counter()
            it_ref, cv_ref = warmstart.solve_residuals_from(s, ref, None, tol=tol)
            t_ref = time.perf_counter() - t0
            aps = [warmstart.features(s, p.Y[:, c].copy())[1] for c in range(K)]
            apc = [np.ascontiguousarray(p.Y[:, c]).copy() for c in range(K)]
            it_ap, cv_ap = warmstart.solve_residuals_from(s, apc, aps, tol=tol)
            lrn = [np.ascontiguousarray(p.Y[:, c]).copy() for c in range(K)]
            t0 = time.perf_counter()
            it_l, cv_l = warmstart.solve_residuals_learned(s, lrn, net, tol=tol)
            t_l = time.perf_counter() - t0
        err = max(float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(lrn, ref))
        rows.append({"seed": int(sd), "N": int(p.N), "G": int(p.info["G"]), "K": int(K),
                     "iters_reference_start": it_ref, "iters_ap_start": it_ap, "iters_learned": it_l,
                     "converged": bool(all(cv_ref) and all(cv_ap) and all(cv_l)),
                     "max_rel_diff_learned_vs_reference": err, "s_reference": t_ref, "s_learned": t_l})
        if log_every and (i + 1) % log_every == 0:
            log(f"[eval] {i + 1}/{len(seeds)} panels")
    return rows


def summarize(rows):
    ref = np.array([i for r in rows for i in r["iters_reference_start"]], float)
    ap = np.array([i for r in rows for i in r["iters_ap_start"]], float)
    ln = np.array([i for r in rows for i in r["iters_learned"]], float)
    return {"columns": int(len(ref)), "mean_iters_reference_start": float(ref.mean()),
            "mean_iters_ap_start": float(ap.mean()), "mean_iters_learned": float(ln.mean()),
            "iters_saved_vs_reference_pct": float(100 * (1 - ln.sum() / ref.sum())),
            "iters_saved_vs_ap_pct": float(100 * (1 - ln.sum() / ap.sum())),
            "all_converged": bool(all(r["converged"] for r in rows)),
            "max_rel_diff_learned_vs_reference": float(max(r["max_rel_diff_learned_vs_reference"] for r in rows))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--panels", type=int, default=2000)
    ap.add_argument("--seed0", type=int, default=10_000)
    ap.add_argument("--n-min", type=float, default=1e4)
    ap.add_argument("--n-max", type=float, default=2e5)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=64)
    ap.add_argument("--eval-seed0", type=int, default=900_000)
    ap.add_argument("--eval-n-max", type=float, default=2e6)
    ap.add_argument("--out", default=warmstart.WEIGHTS_PATH)
    ap.add_argument("--report", default="")
    args = ap.parse_args()

    t0 = time.perf_counter()
    X, t = warmstart.build_training_set(args.panels, args.seed0, args.n_min, args.n_max, log=log)
    t_data = time.perf_counter() - t0
    log(f"[train] {len(t)} rows from {args.panels} panels in {t_data:.1f} s; target rms {np.sqrt(np.mean(t * t)):.4f}")
    net = warmstart.WarmStartNet(seed=1)
    t0 = time.perf_counter()
    losses = net.fit(X, t, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=2)
    t_fit = time.perf_counter() - t0
    log(f"[train] losses {np.round(losses, 5).tolist()} ({t_fit:.1f} s)")
    net.save(args.out, train_panels=args.panels, n_min=args.n_min, n_max=args.n_max, epochs=args.epochs,
             losses=losses)
    rows = evaluate(net, list(range(args.eval_seed0, args.eval_seed0 + args.eval)), args.eval_n_max)
    summ = summarize(rows)
    summ.update({"train_rows": int(len(t)), "train_panels": args.panels, "train_seconds": t_fit,
                 "data_seconds": t_data, "loss_per_epoch": losses.tolist(),
                 "target_mean_sq_half": float(0.5 * np.mean(t * t))})
    log(json.dumps(summ))
    if args.report:
        os.makedirs(os.path.dirname(os.path.abspath(args.report)), exist_ok=True)
        json.dump({"summary": summ, "panels": rows}, open(args.report, "w"), indent=1)
    print(json.dumps(summ))


if __name__ == "__main__":
    main()
