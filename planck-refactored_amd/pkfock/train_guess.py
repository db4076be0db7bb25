"""Train and evaluate the learned initial guess for water clusters (6-31G*).

    python -m pkfock.train_guess --train 512 --n-max 20 --eval 8 --eval-n 48 \
        --out model.npz --report report.json

Training set: seeded water clusters (pkfock.molecules.water_cluster, seeds
100000 + i, 2..n_max molecules), each converged by this engine's device RHF
(core guess, Schwarz 1e-12, tol_energy 1e-9, tol_density 1e-7); the converged
densities are cut into frame-local blocks and the two networks are trained on
the device (pkfock.guess.LearnedGuess.fit).  Held-out set: seeds 900000 + i
with eval_n molecules.  For each held-out cluster the RHF runs from the core
guess and from the learned density (with the fallback rule of
pkfock.guess.scf_with_guess) to tol_energy 1e-8, tol_density 1e-6; iterations,
wall times (the learned run includes the network, the feature construction and
the two one-iteration probes of the fallback rule) and the energy difference
are reported.
"""
import argparse
import json
import time

import numpy as np

from .guess import LearnedGuess, scf_with_guess
from .host import Molecule
from .molecules import water_cluster

SCF_TRAIN = dict(tol_eri=1e-12, tol_energy=1e-9, tol_density=1e-7)
SCF_EVAL = dict(tol_eri=1e-12, tol_energy=1e-8, tol_density=1e-6)


def training_set(n, n_min, n_max, log):
    out = []
    for i in range(n):
        Z, X = water_cluster(100000 + i, n_min=n_min, n_max=n_max)
        r = Molecule(Z, X, "6-31g*").rhf(**SCF_TRAIN)
        if r["rc"] != 0:
            log("train cluster %d (%d waters) did not converge; skipped" % (i, len(Z) // 3))
            continue
        out.append((X, r["P"]))
        if i % 32 == 0:
            log("train cluster %d: %d waters, %d iterations" % (i, len(Z) // 3, r["iterations"]))
    return out


def evaluate(model, n_eval, n_waters, log):
    rows = []
    for i in range(n_eval):
        Z, X = water_cluster(900000 + i, n=n_waters)
        mol = Molecule(Z, X, "6-31g*")
        t = time.perf_counter()
        core = mol.rhf(**SCF_EVAL)
        t_core = time.perf_counter() - t
        t = time.perf_counter()
        lrn, used = scf_with_guess(mol, model, **SCF_EVAL)
        t_lrn = time.perf_counter() - t
        row = {"seed": 900000 + i, "nbasis": mol.nbasis, "core_iterations": core["iterations"],
               "learned_iterations": lrn["iterations"], "guess_used": used, "core_wall_s": t_core,
               "learned_wall_s": t_lrn, "dE": lrn["energy"] - core["energy"], "energy": core["energy"],
               "converged": bool(core["rc"] == 0 and lrn["rc"] == 0)}
        rows.append(row)
        log("eval %d: nb=%d core %d it %.2fs, learned (%s) %d it %.2fs, dE %.2e" % (
            i, mol.nbasis, core["iterations"], t_core, used, lrn["iterations"], t_lrn, row["dE"]))
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=512)
    ap.add_argument("--n-min", type=int, default=2)
    ap.add_argument("--n-max", type=int, default=20)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--eval-n", type=int, default=48)
    ap.add_argument("--out", default="guess_model.npz")
    ap.add_argument("--report", default="guess_report.json")
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    import torch
    torch.cuda.init()  # torch (device buffers of the networks) first, then the engine
    t0 = time.perf_counter()
    data = training_set(args.train, args.n_min, args.n_max, log)
    t_data = time.perf_counter() - t0
    model = LearnedGuess(hidden=128, seed=0)
    t0 = time.perf_counter()
    hist = model.fit(data, epochs=args.epochs, batch=512, lr=2e-3, log=log)
    t_fit = time.perf_counter() - t0
    np.savez(args.out, **model.state())
    rows = evaluate(model, args.eval, args.eval_n, log)
    ic = np.array([r["core_iterations"] for r in rows], float)
    il = np.array([r["learned_iterations"] for r in rows], float)
    rep = {"train_clusters": len(data), "train_sizes": [args.n_min, args.n_max], "epochs": args.epochs,
           "train_data_s": t_data, "fit_s": t_fit,
           "final_loss": {"intra": [h[2] for h in hist if h[0] == "xi"][-1], "inter": [h[2] for h in hist if h[0] == "xe"][-1]},
           "eval_clusters": len(rows), "eval_waters": args.eval_n,
           "mean_core_iterations": ic.mean(), "mean_learned_iterations": il.mean(),
           "iteration_reduction": 1.0 - il.sum() / ic.sum(),
           "mean_core_wall_s": float(np.mean([r["core_wall_s"] for r in rows])),
           "mean_learned_wall_s": float(np.mean([r["learned_wall_s"] for r in rows])),
           "max_abs_dE": float(np.max(np.abs([r["dE"] for r in rows]))), "fallbacks": sum(r["guess_used"] != "learned" for r in rows),
           "rows": rows}
    json.dump(rep, open(args.report, "w"), indent=1)
    log("iterations %.1f -> %.1f (%.0f%% fewer), max |dE| %.2e" % (ic.mean(), il.mean(), 100 * rep["iteration_reduction"], rep["max_abs_dE"]))


if __name__ == "__main__":
    main()
