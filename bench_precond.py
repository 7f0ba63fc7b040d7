"""Time-to-misfit of the learned preconditioner against the plain method.

Workload (the learned-preconditioner configuration over the 134x384 model
family): seeded synthetic models of 12 dipping layers, Vp 1500-4700 m/s with
a +/-5% smooth perturbation (correlation 15 cells), 24 m cells, 32-cell C-PML
(224x448 padded), nt 2000, dt 2 ms, 30 shots, 384 receivers, 8 Hz; observed
data from the true model, inversion of cp from the Gaussian-smoothed model
with cs and rho fixed (as the reference's plain inversion does).  The top 10
rows are elastic, not water: with a Vs = 0 layer the operator's adjoint (the
reference's scheme) grows exponentially over 2000 steps and its gradient is
not a descent direction, so no optimiser can use it.

1. Plain inversions (L-BFGS-B, bounds 1500-5500, ftol = gtol = 1e-6, the
   reference's settings) on the training seeds log (gradient, illumination,
   accepted update) pairs in memory.
2. The network (HIP kernels) trains on those pairs on the GPU.
   With ``--teacher control`` the pairs come instead from the L-BFGS driver
   with the illumination-scaled H0 (it does not stall on the approximate
   gradient, so it logs more and later updates).
3. On held-out seeds: the plain method runs to its end (maxiter); the learned
   driver then runs until it reaches the plain method's final misfit.
   value = median over held-out seeds of (learned wall time / plain wall time);
   lower is better, the target is <= 0.5.  ``--driver lbfgs`` (default) is
   L-BFGS with the network as the initial inverse Hessian
   (``precond.learned_lbfgs``); ``--driver descent`` the network's direction
   with a line search (``precond.learned_descent``); ``--driver diag`` the
   plain method's L-BFGS-B on variables scaled by the diagonal the network
   implies at the starting model (``precond.diagonal_scale``).  The same
   driver with the illumination scaling and no network runs as a control.

``--adjoint exact`` gives every objective the exact gradient (the transpose
of the forward step) instead of the reference's.  ``--init-net`` continues
training from saved parameters and ``--eval-first`` / ``--time-cap`` split
and bound the evaluation, so a full-size run fits several shorter sessions.

Prints ONE JSON line; progress goes to stderr.  The defaults are the
configuration that met the target (exact adjoint, ``--driver diag`` with the
network's diagonal times the illumination one, clip [0.1, 10]), sized to
finish in about 15 minutes on one MI355X.  The recorded run trained in two
stages (``--train-models 10 --seed0 100 --teacher-maxiter 40 --train-only
--save-net A``, then ``--train-models 10 --seed0 110 --init-net A --save-net
B``) and evaluated with ``--load-net B --eval-models 5 --maxiter 60``.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fwiflow.jl_amd"))

import numpy as np  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--train-models", type=int, default=10)
    p.add_argument("--eval-models", type=int, default=3)
    p.add_argument("--maxiter", type=int, default=60)
    p.add_argument("--nt", type=int, default=2000)
    p.add_argument("--epochs", type=int, default=8)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--step0", type=float, default=20.0)
    p.add_argument("--seed0", type=int, default=100)
    p.add_argument("--driver", choices=["lbfgs", "descent", "diag"], default="diag")
    p.add_argument("--teacher", choices=["plain", "control"], default="control",
                   help="optimiser whose accepted updates are the training pairs")
    p.add_argument("--teacher-maxiter", type=int, default=None,
                   help="iterations of each training inversion (default: --maxiter)")
    p.add_argument("--save-net", default=None, help="write the trained parameters (.npy)")
    p.add_argument("--load-net", default=None,
                   help="skip training and evaluate the parameters in this .npy file")
    p.add_argument("--init-net", default=None,
                   help="start training from these parameters (.npy) instead of a fresh "
                        "network: training in stages over several runs")
    p.add_argument("--train-only", action="store_true", help="stop after training")
    p.add_argument("--time-cap", type=float, default=None,
                   help="stop the learned and control runs at this multiple of the plain "
                        "method's wall time (their ratio is then censored at the cap)")
    p.add_argument("--eval-seed0", type=int, default=None,
                   help="seed of held-out model 0 (default seed0 + 1000); the training seeds "
                        "serve as a validation set for the diagonal's settings")
    p.add_argument("--diag-clip", type=float, nargs=2, default=(0.1, 10.0))
    p.add_argument("--diag-smooth", type=float, default=3.0)
    p.add_argument("--diag-illum-power", type=float, default=1.0)
    p.add_argument("--no-control", action="store_true", help="skip the control runs")
    p.add_argument("--eval-first", type=int, default=0,
                   help="index of the first held-out model (seed seed0 + 1000 + index): "
                        "evaluation split over several runs")
    p.add_argument("--adjoint", choices=["reference", "exact"], default="exact",
                   help="gradient of every objective: the reference's adjoint or the exact "
                        "one (the transpose of the forward step)")
    return p.parse_args()


def setup(seed, nt, ws, adjoint="reference"):
    import fwiflow
    from fwiflow import models

    c = models.case2(seed=seed, water=False)
    stf = models.sourceGene(c.f0, nt, c.dt) if nt != c.nSteps else c.stf
    fwi = fwiflow.make_fwi(c.nz, c.nx, c.dz, c.dx, nt, c.dt, c.src_z, c.src_x, c.rec_z, c.rec_x,
                           nPml=c.nPml, f0=c.f0, WORKSPACE=ws)
    fwiflow.compute_observation(fwi, c.cp, c.cs, c.rho, stf)
    obj = fwiflow.PlanObjective(fwi, c.cs, c.rho, stf, cp_ref=c.cp0, adjoint=adjoint)
    x0 = fwiflow.padding(c.cp0, fwi.nPml, fwi.nPad)
    return fwi, obj, x0


def main():
    a = parse()
    import fwiflow
    from fwiflow import inversion, precond

    t_all = time.perf_counter()
    mask = None
    samples = []
    plain_train = []
    train_models = 0 if a.load_net else a.train_models
    for k in range(train_models):
        seed = a.seed0 + k
        ws = tempfile.mkdtemp(prefix=f"fwiflow_pc_{seed}_")
        fwi, obj, x0 = setup(seed, a.nt, ws, a.adjoint)
        mask = fwi.mask

        def on_pair(g, extras, update, J):
            x = precond.features(extras["gradients"], extras["illumination"], mask)
            samples.append((x, precond.target_from_update(update, mask)))

        tmax = a.teacher_maxiter or a.maxiter
        if a.teacher == "plain":
            r = inversion.lbfgs(obj, x0, maxiter=tmax, on_pair=on_pair)
        else:
            r = precond.learned_lbfgs(obj, x0, mask, net=None, maxiter=tmax, on_pair=on_pair)
        obj.close()
        plain_train.append((seed, r["iterations"], r["evaluations"], r["seconds"]))
        log(f"train model {seed}: {r['iterations']} iterations, {r['evaluations']} evaluations, "
            f"{r['seconds']:.1f} s, misfit {r['history'][0][1] if r['history'] else 0:.4g} -> "
            f"{r['misfit']:.4g}; pairs {len(samples)}")
    if a.load_net:
        fwi0, obj0, _ = setup(a.seed0, a.nt, tempfile.mkdtemp(prefix="fwiflow_pc_mask_"))
        mask = fwi0.mask
        obj0.close()
    H, W = mask.shape
    net = precond.CNN(H, W, seed=0)
    t0 = time.perf_counter()
    if a.load_net:
        net.set_params(np.load(a.load_net).astype(np.float32))
        hist = [float("nan")]
        log(f"loaded network parameters from {a.load_net}")
    else:
        if a.init_net:
            net.set_params(np.load(a.init_net).astype(np.float32))
            log(f"training continues from {a.init_net}")
        hist = precond.train(net, samples, mask.astype(np.float32), epochs=a.epochs, lr=a.lr,
                             log=log)
        log(f"trained on {len(samples)} pairs in {time.perf_counter() - t0:.1f} s, "
            f"loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    t_train = time.perf_counter() - t0
    if a.save_net:
        np.save(a.save_net, net.get_params())
    if a.train_only:
        print(json.dumps({"training": {"pairs": len(samples), "seconds": round(t_train, 1),
                                       "loss": hist, "models": train_models},
                          "plain_training_runs": [{"seed": s_, "iterations": i, "evaluations": e,
                                                   "seconds": round(t, 1)}
                                                  for s_, i, e, t in plain_train]}), flush=True)
        net.close()
        return
    evals = []
    for k in range(a.eval_first, a.eval_first + a.eval_models):
        seed = (a.seed0 + 1000 if a.eval_seed0 is None else a.eval_seed0) + k
        ws = tempfile.mkdtemp(prefix=f"fwiflow_pc_{seed}_")
        fwi, obj, x0 = setup(seed, a.nt, ws, a.adjoint)
        J0 = obj(x0)[0]
        rp = inversion.lbfgs(obj, x0, maxiter=a.maxiter)
        log(f"eval {seed}: plain {rp['misfit']:.4g} in {rp['seconds']:.1f} s, "
            f"{rp['evaluations']} evaluations")

        def progress(k, J, used, alpha, nev):
            if k % 5 == 0:
                log(f"  iteration {k}: misfit {J:.4g} ({used}, alpha {alpha:.3g}, {nev} evaluations)")
        cap = a.time_cap * rp["seconds"] if a.time_cap else None

        def diag(net_):
            # one evaluation at the start gives the scaling, then the plain
            # method's L-BFGS-B on the scaled variables (both timed)
            t0 = time.perf_counter()
            _, g0, ex0 = obj(x0)
            s = precond.diagonal_scale(g0, ex0, fwi.mask, net_, smooth=a.diag_smooth,
                                       clip=tuple(a.diag_clip),
                                       illum_power=a.diag_illum_power if net_ else 0.0)
            r = inversion.lbfgs(obj, x0, maxiter=2 * a.maxiter, target_misfit=rp["misfit"],
                                scale=s)
            r["seconds"] = time.perf_counter() - t0
            r["evaluations"] += 1
            r["history"] = [(t, J, "diag") for t, J in r["history"]]
            return r
        if a.driver == "diag":
            rl = diag(net)
        elif a.driver == "lbfgs":
            rl = precond.learned_lbfgs(obj, x0, fwi.mask, net=net, maxiter=2 * a.maxiter,
                                       target_misfit=rp["misfit"], on_iter=progress,
                                       max_seconds=cap)
        else:
            rl = precond.learned_descent(obj, x0, net, fwi.mask, maxiter=2 * a.maxiter,
                                         target_misfit=rp["misfit"], step0=a.step0)
        log(f"eval {seed}: plain {rp['misfit']:.4g} in {rp['seconds']:.1f} s; "
            f"learned {rl['misfit']:.4g} in {rl['seconds']:.1f} s")
        if a.no_control:
            rc = {"misfit": float("inf"), "seconds": float("nan")}
        elif a.driver == "diag":
            rc = diag(None)
        else:
            rc = precond.learned_lbfgs(obj, x0, fwi.mask, net=None, maxiter=2 * a.maxiter,
                                       target_misfit=rp["misfit"], on_iter=progress,
                                       max_seconds=cap)
        log(f"eval {seed}: control {rc['misfit']:.4g} in {rc['seconds']:.1f} s")
        obj.close()
        reached = rl["misfit"] <= rp["misfit"]
        used = [h[2] for h in rl["history"]]
        e = {"seed": seed, "misfit_start": J0, "plain_misfit": rp["misfit"],
             "plain_seconds": round(rp["seconds"], 2), "plain_iterations": rp["iterations"],
             "plain_evaluations": rp["evaluations"], "learned_misfit": rl["misfit"],
             "learned_seconds": round(rl["seconds"], 2), "learned_iterations": rl["iterations"],
             "learned_evaluations": rl["evaluations"], "reached": bool(reached),
             "learned_steps": len(used) - used.count("fallback"),
             "fallback_steps": used.count("fallback"),
             "ratio": (rl["seconds"] / rp["seconds"]) if reached else None,
             "control_misfit": rc["misfit"], "control_seconds": round(rc["seconds"], 2),
             "control_ratio": (rc["seconds"] / rp["seconds"])
             if rc["misfit"] <= rp["misfit"] else None}
        evals.append(e)
        log("eval", json.dumps(e))
    ratios = [e["ratio"] for e in evals]
    value = float(np.median([r if r is not None else np.inf for r in ratios])) if evals else None
    out = {
        "metric": "time-to-misfit, learned preconditioner / plain L-BFGS (134x384 model family, "
                  "elastic top layer)",
        "value": value if value is None or np.isfinite(value) else None,
        "unit": "ratio of wall times", "higher_is_better": False,
        "target": 0.5, "n_gpus": 1, "data": "synthetic (seeded layered models)",
        "control_value": (lambda v: v if np.isfinite(v) else None)(float(np.median(
            [e["control_ratio"] if e["control_ratio"] is not None else np.inf
             for e in evals]))) if evals else None,
        "config": {"driver": a.driver, "teacher": a.teacher, "adjoint": a.adjoint,
                   "grid": [int(H), int(W)],
                   "nt": a.nt, "shots": 30,
                   "train_models":
                   a.train_models, "eval_models": a.eval_models, "eval_first": a.eval_first,
                   "maxiter": a.maxiter, "time_cap": a.time_cap,
                   "diag": {"clip": list(a.diag_clip), "smooth": a.diag_smooth,
                            "illum_power": a.diag_illum_power},
                   "pairs": len(samples), "epochs": a.epochs, "lr": a.lr},
        "training": {"seconds": round(t_train, 1), "loss_first_epoch": hist[0],
                     "loss_last_epoch": hist[-1]},
        "plain_training_runs": [{"seed": s, "iterations": i, "evaluations": e,
                                 "seconds": round(t, 1)} for s, i, e, t in plain_train],
        "eval": evals, "total_seconds": round(time.perf_counter() - t_all, 1),
    }
    net.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
