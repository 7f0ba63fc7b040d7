"""Plain FWI driver: L-BFGS-B with bounds, forward illumination, pair logging.

The reference's plain inversion is scipy's L-BFGS-B behind ADCME's
``ScipyOptimizerInterface`` on the padded P velocity, with bounds
(1500, 5500) m/s and options maxiter = 100, ftol = gtol = 1e-6
(docs/codes/src_fwi_coupled/main_mar_fwi.jl:184-187; ``BFGS!`` in
test/TestFWI.jl:74).  Its loss is ``compute_misfit`` with the masked blend
cp * mask + cp_ref * (1 - mask) (src/FWI.jl:156-189); its callbacks log the
loss of every evaluation and (cp, gradient) of every iteration
(main_mar_fwi.jl:127-171).

This driver runs the same optimiser over the device-resident plan:

* ``PlanObjective`` evaluates misfit and gradient with ONE calc_id 1 run (the
  plan keeps the misfit its gradient run computes; the TF graph needs a
  calc_id 0 run plus a calc_id 1 run), with the observed data resident in HBM
  and, optionally, the forward illumination summed in the same run.
* ``lbfgs`` is scipy's L-BFGS-B with the reference's defaults.  Every
  evaluation is logged (misfit, wall time: the time-to-misfit curve) and every
  accepted iteration writes one (gradient, illumination, accepted update)
  pair, the training sample of a learned preconditioner.

The objective is any callable ``f(x) -> (misfit, gradient, extras)``, so the
driver itself runs without a GPU.
"""
from __future__ import annotations

import json
import os
import time
from collections import OrderedDict

import numpy as np
from scipy.optimize import minimize

from .fwi import _on_grid, _stf_rows, try_pad
from .ops import Plan
from .utils import moduli_to_velocity_grad, velocity_to_moduli

# reference defaults (main_mar_fwi.jl:184-187)
CP_BOUNDS = (1500.0, 5500.0)
MAXITER = 100
FTOL = 1e-6
GTOL = 1e-6


class PlanObjective:
    """Misfit and dJ/dcp of the masked FWI loss on one device-resident plan.

    ``fwi``: geometry from ``make_fwi``; ``cs``, ``rho``: fixed models;
    ``cp_ref``/``cs_ref``/``rho_ref``: the models blended in outside the mask
    (``is_masked = False`` of compute_misfit); ``obs``: observed data as
    written by ``compute_observation``'s files, either a list of [nrec][nSteps]
    float32 arrays per shot or None to load ``Shot{id}.bin`` from the
    workspace.  shot_ids are 1-based as in the reference's Julia API.
    ``adjoint``: "reference" (the operator's gradient as cufd computes it) or
    "exact" (the transpose of the forward step: the derivative of the
    discrete misfit).
    """

    def __init__(self, fwi, cs, rho, stf_array, cp_ref, cs_ref=None, rho_ref=None,
                 shot_ids=None, obs=None, gpu_id=0, illumination=True, adjoint="reference"):
        self.fwi = fwi
        if shot_ids is None:
            shot_ids = np.arange(1, len(fwi.ind_src_x) + 1)
        self.ids = np.asarray(shot_ids) - 1
        stf = _stf_rows(fwi, stf_array)
        self.cs = try_pad(fwi, cs)
        self.rho = try_pad(fwi, rho)
        self.cp_ref = try_pad(fwi, cp_ref)
        self.cs_ref = self.cs if cs_ref is None else try_pad(fwi, cs_ref)
        self.rho_ref = self.rho if rho_ref is None else try_pad(fwi, rho_ref)
        _on_grid(fwi, self.cs, self.rho, self.cp_ref, self.cs_ref, self.rho_ref)
        self.mask = fwi.mask
        self.plan = Plan(fwi.para_path, self.ids, gpu_id)
        self.plan.set_adjoint(adjoint)
        self.plan.set_stf(stf)
        if obs is None:
            self.plan.load_obs()
        else:
            self.plan.set_obs(obs)
        self.illumination = bool(illumination)
        self.plan.set_illumination(self.illumination)
        self.shape = (fwi.nz_pad, fwi.nx_pad)

    def close(self):
        self.plan.close()

    def models(self, cp):
        """Masked blend of FWI.jl:172-177."""
        m, mn = self.mask, 1.0 - self.mask
        return (cp * m + self.cp_ref * mn, self.cs * m + self.cs_ref * mn,
                self.rho * m + self.rho_ref * mn)

    def __call__(self, x):
        cp = np.asarray(x, dtype=np.float64).reshape(self.shape)
        cpm, csm, rhom = self.models(cp)
        lam, mu = velocity_to_moduli(cpm, csm, rhom)
        P = self.plan
        P.set_model(lam, mu, rhom)
        P.run(1)
        J = P.misfit()
        gL, gM, gD, _ = P.gradients()
        g_cp, g_cs, g_rho = (g * self.mask for g in
                             moduli_to_velocity_grad(cpm, csm, rhom, gL, gM, gD))
        extras = {"gradients": (g_cp, g_cs, g_rho)}
        if self.illumination:
            extras["illumination"] = P.illumination()
        return J, g_cp.ravel(), extras


class _EvalCache:
    """The last `size` evaluations by point (least recently stored first out);
    storing a point again refreshes it."""

    def __init__(self, size=8):
        self.size = size
        self.d = OrderedDict()

    def put(self, x, value):
        key = np.asarray(x).tobytes()
        self.d[key] = value
        self.d.move_to_end(key)
        while len(self.d) > self.size:
            self.d.popitem(last=False)

    def get(self, x):
        return self.d.get(np.asarray(x).tobytes())


class _Log:
    def __init__(self, log_dir):
        self.dir = log_dir
        if log_dir:
            os.makedirs(os.path.join(log_dir, "pairs"), exist_ok=True)

    def line(self, name, rec):
        if self.dir:
            with open(os.path.join(self.dir, name), "a") as f:
                f.write(json.dumps(rec) + "\n")

    def pair(self, k, shape, grad, extras, update, misfit):
        if not self.dir:
            return None
        path = os.path.join(self.dir, "pairs", f"pair_{k:04d}.npz")
        arrays = {"gradient": grad.reshape(shape).astype(np.float32),
                  "update": update.reshape(shape).astype(np.float32),
                  "misfit": np.float64(misfit)}
        if extras.get("illumination") is not None:
            arrays["illumination"] = np.asarray(extras["illumination"],
                                                dtype=np.float32).reshape(shape)
        if extras.get("gradients") is not None:
            for name, gq in zip(("gradient_cp", "gradient_cs", "gradient_rho"),
                                extras["gradients"]):
                arrays[name] = np.asarray(gq, dtype=np.float32).reshape(shape)
        np.savez(path, **arrays)
        return path


def lbfgs(objective, x0, bounds=CP_BOUNDS, maxiter=MAXITER, ftol=FTOL, gtol=GTOL,
          log_dir=None, shape=None, target_misfit=None, on_pair=None, first_step_rms=10.0,
          scale=None):
    """Minimise ``objective`` with L-BFGS-B from ``x0``.

    ``objective(x) -> (misfit, gradient, extras)`` with ``extras`` optionally
    holding ``"illumination"``.  ``bounds``: a (low, high) pair for every
    variable.  With ``log_dir``: ``evals.jsonl`` (one line per evaluation:
    misfit and seconds since the start), ``iters.jsonl`` (one per accepted
    iteration) and ``pairs/pair_NNNN.npz`` (gradient and illumination at the
    iterate, and the accepted update to the next iterate); ``on_pair(gradient,
    extras, update, misfit)`` receives the same pairs in memory.
    ``target_misfit`` stops the run once an accepted iterate reaches it
    (time-to-misfit runs).

    ``first_step_rms``: L-BFGS-B's first trial step has unit 2-norm in its
    variables; with cp in m/s over ~1e5 cells that is ~0.005 m/s per cell,
    below what an FP32 forward model's misfit resolves, and its line search
    then fails.  The variables are therefore scaled so that a unit step is
    `first_step_rms` m/s RMS per cell (gtol then applies to the scaled
    gradient); None runs on cp itself as the reference's
    ScipyOptimizerInterface does.  ``scale``: optional positive factors, one
    per variable (a diagonal preconditioner: a unit step moves variable i by
    scale[i] times as much).

    Returns a dict: x, misfit, iterations, evaluations, seconds, history
    (per-iteration (seconds, misfit)), scipy's message.
    """
    x0 = np.asarray(x0, dtype=np.float64)
    shape = x0.shape if shape is None else shape
    x0 = x0.ravel()
    k = 1.0 if first_step_rms is None else 1.0 / (float(first_step_rms) * np.sqrt(x0.size))
    if scale is not None:
        k = k / np.asarray(scale, dtype=np.float64).ravel()
    lo, hi = bounds
    y0 = x0 * k
    to_x = lambda y: np.clip(np.asarray(y) / k, lo, hi)  # noqa: E731
    x0 = to_x(y0)  # the point scipy evaluates first, bit for bit
    log = _Log(log_dir)
    t0 = time.perf_counter()
    cache = _EvalCache()  # (misfit, gradient, extras) of recent evaluations
    state = {"evals": 0, "iters": 0, "x": x0.copy(), "history": []}

    def fun(x):
        J, g, extras = objective(x)
        g = np.asarray(g, dtype=np.float64).ravel()
        state["evals"] += 1
        cache.put(x, (float(J), g.copy(), extras))
        log.line("evals.jsonl", {"eval": state["evals"], "misfit": float(J),
                                 "seconds": time.perf_counter() - t0})
        return float(J), g

    class _Stop(Exception):
        pass

    def callback(xk):
        prev = state["x"]
        ent = cache.get(prev)
        if ent is None:  # the previous iterate was evicted: evaluate it again
            fun(prev)
            ent = cache.get(prev)
        Jp, gp, extras = ent
        cur = cache.get(xk)
        state["iters"] += 1
        k = state["iters"]
        t = time.perf_counter() - t0
        Jk = cur[0] if cur else float("nan")
        log.pair(k, shape, gp, extras, np.asarray(xk) - prev, Jp)
        if on_pair is not None:
            on_pair(gp, extras, np.asarray(xk) - prev, Jp)
        log.line("iters.jsonl", {"iter": k, "misfit": Jk, "seconds": t,
                                 "evals": state["evals"]})
        state["history"].append((t, Jk))
        state["x"] = np.array(xk, dtype=np.float64)
        if target_misfit is not None and cur and cur[0] <= target_misfit:
            raise _Stop()

    def fun_y(y):
        J, g = fun(to_x(y))
        return J, g / k

    def callback_y(yk):
        callback(to_x(yk))

    res_x, msg = None, ""
    try:
        r = minimize(fun_y, y0, jac=True, method="L-BFGS-B",
                     bounds=list(zip(np.broadcast_to(lo * k, x0.shape),
                                     np.broadcast_to(hi * k, x0.shape))),
                     callback=callback_y,
                     options={"maxiter": maxiter, "ftol": ftol, "gtol": gtol})
        res_x, msg = to_x(r.x), str(r.message)
    except _Stop:
        res_x, msg = state["x"], "target misfit reached"
    final = cache.get(res_x)
    J = final[0] if final else fun(res_x)[0]
    return {"x": res_x.reshape(shape), "misfit": J, "iterations": state["iters"],
            "evaluations": state["evals"], "seconds": time.perf_counter() - t0,
            "history": state["history"], "message": msg}
