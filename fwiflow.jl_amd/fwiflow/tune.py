"""Launch-shape autotuner: tile width and shots per launch, measured and cached.

The engine's automatic shape (64 x 4 workgroups; shots per launch from the
Infinity-Cache rule, round(224 MB / bytes per shot)) is a model; this module
measures the neighbourhood of it on the caller's own grid and shot count and
keeps the fastest, per (grid, C-PML, shot count), in a JSON cache
(``FWIFLOW_TUNE_CACHE`` or ``~/.cache/fwiflow/launch.json``).

A candidate is timed on the caller's model over a shortened record
(``steps`` time steps; every step of a run costs the same, so the ranking
carries over to the full record): one gradient run (calc_id 1) to warm up,
then the best of ``repeats`` runs.  Shapes only change how the shots are
grouped into launches, never the arithmetic of a cell, so every candidate
gives the same data and, up to the order of the gradient sums over a
launch's shots, the same gradients (tests/test_gpu.py checks each switch).
"""
from __future__ import annotations

import json
import os
import tempfile
import time

import numpy as np

from .ops import Plan

ARCH = "gfx950"


def cache_path(path=None):
    return path or os.environ.get("FWIFLOW_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "fwiflow", "launch.json")


def cache_key(para_fname, nshots):
    with open(para_fname) as f:
        p = json.loads(f.read())
    return f"{ARCH}/{p['nz']}x{p['nx']}/pml{p['nPoints_pml']}/shots{int(nshots)}"


def load_cache(path=None):
    path = cache_path(path)
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def store(key, entry, path=None):
    path = cache_path(path)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    c = load_cache(path)
    c[key] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(c, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def lookup(para_fname, nshots, path=None):
    """The cached launch shape for this grid and shot count, or None."""
    return load_cache(path).get(cache_key(para_fname, nshots))


def apply_tuned(plan, para_fname, nshots, path=None):
    """Set the cached shape on ``plan``; returns it (None: nothing cached,
    the plan keeps the automatic shape)."""
    e = lookup(para_fname, nshots, path)
    if e is not None:
        plan.set_launch(e["tile_x"], e["shots_fwd"], e["shots_bwd"])
    return e


def _short_para(para_fname, steps, ws):
    with open(para_fname) as f:
        p = json.loads(f.read())
    n = min(int(steps), int(p["nSteps"]))
    p["nSteps"] = n
    p["data_dir_name"] = os.path.join(ws, "Data")
    p.pop("scratch_dir_name", None)
    os.makedirs(p["data_dir_name"], exist_ok=True)
    out = os.path.join(ws, "para_file.json")
    with open(out, "w") as f:
        f.write(json.dumps(p))
    return out, n


def autotune(para_fname, shot_ids, lam, mu, den, stf, steps=200, tiles=(4, 2), repeats=2,
             gpu_id=0, path=None, log=None):
    """Time launch shapes around the automatic one on ``steps`` time steps of
    the caller's problem; cache and return the fastest as a dict
    (tile_x, shots_fwd, shots_bwd, seconds per gradient run of ``steps``,
    and every candidate's time)."""
    say = log or (lambda *a: None)
    ids = list(np.atleast_1d(shot_ids))
    stf = np.atleast_2d(np.asarray(stf, dtype=np.float64))
    with tempfile.TemporaryDirectory(prefix="fwiflow_tune_") as ws:
        para, n = _short_para(para_fname, steps, ws)
        with Plan(para, ids, gpu_id) as P:
            P.set_model(lam, mu, den)
            P.set_stf(np.ascontiguousarray(stf[:, :n]))
            P.run(2)
            P.set_obs(P.data())
            P.set_launch(4, -1, -1)
            P.run(1)
            f0, b0 = P.shots_per_launch()
            ns = len(ids)
            near = lambda k: sorted({max(1, min(ns, k + d)) for d in (-1, 0, 1)})  # noqa: E731
            cands = [(t, f, b) for t in tiles for f in near(f0) for b in near(b0)]
            times = {}
            for t, f, b in cands:
                P.set_launch(t, f, b)
                P.run(1)
                best = np.inf
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    P.run(1)
                    best = min(best, time.perf_counter() - t0)
                times[(t, f, b)] = best
                say(f"tune tile {t} shots {f}/{b}: {best * 1e3:.1f} ms")
    (t, f, b), sec = min(times.items(), key=lambda kv: kv[1])
    entry = {"tile_x": t, "shots_fwd": f, "shots_bwd": b, "seconds": sec, "steps": n,
             "automatic": [4, f0, b0], "automatic_seconds": times.get((4, f0, b0)),
             "candidates": {f"{k[0]}/{k[1]}/{k[2]}": v for k, v in times.items()}}
    store(cache_key(para_fname, len(ids)), entry, path)
    return entry
