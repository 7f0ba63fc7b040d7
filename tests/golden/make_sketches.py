"""Writes the oracle's results for the large GPU parity cases as sketches.

The 500x1700 gradient (2 shots, 1500 steps), the central finite-difference
check on that grid, the time-lapse surveys on the 200x600, 300x900 and
450x1350 grids at nt 3000 and the 134x384 surveys at nt 2000 (water layer
and elastic top, 3 shots) take the CPU oracle minutes per case, too long for
the GPU test run; this script runs them once (FP32 reference arithmetic and
FP64) and stores, per case, 64-probe sketches of the seismograms and of the
lambda/mu/rho gradients (tests/sketch.py), the misfits, the source-function
gradients and the finite-difference values, in tests/golden/sketch_*.npz.

    python tests/golden/make_sketches.py [c3|fd|c4|c2 ...]

Inputs are the seeded models of fwiflow.models; the oracle's observed data
come from its own FP32 forward run (calc_id 2).
"""
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "fwiflow.jl_amd"), os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

from fwiflow import models  # noqa: E402
from oracle import oracle  # noqa: E402
from sketch import sketch, smooth_direction  # noqa: E402

C3_NT, C3_IDS = 1500, [0, 40]
C4_NT = 3000
C4_GRIDS = [(200, 600), (300, 900), (450, 1350)]
C2_IDS = [0, 14, 29]
FD_H, FD_DIRS = 1e-3, 5


def log(*a):
    print(*a, flush=True)


def c4_ids(nz, nx, nshots=16):
    """The shot nearest above the monitor ellipse (same draws as models.case4)
    and the first shot."""
    rng = np.random.default_rng(4)
    rng.uniform()
    cx = rng.uniform(0.35, 0.65) * nx
    src_x = np.round(np.linspace(5, nx - 6, nshots)).astype(int)
    return [int(np.argmin(np.abs(src_x - cx))), 0]


def grad_case(case, ids, ws, gradient=True):
    """Oracle seismograms, misfit and (unless gradient=False) gradients at the
    starting model, FP32 and FP64."""
    para, stf = models.write_case(case, ws)
    lam, mu, den = models.moduli(case, "true")
    out = {"ids": np.array(ids)}
    for prec in (32, 64):
        t0 = time.time()
        o = oracle.cufd(lam, mu, den, stf, 2, ids, para, precision=prec)[5]
        out[f"seis{prec}"] = sketch(np.concatenate([s.ravel() for s in o["syn"]]))
        out[f"seis{prec}_norm"] = np.linalg.norm(np.concatenate([s.ravel() for s in o["syn"]]))
        if prec == 32:  # observed data for both precisions: the FP32 run's Shot files
            syn32 = np.concatenate([s.ravel() for s in o["syn"]])
        log(case.name, "forward", prec, f"{time.time() - t0:.1f}s")
    oracle.cufd(lam, mu, den, stf, 2, ids, para, precision=32)  # rewrite the FP32 Shot files
    lam0, mu0, den0 = models.moduli(case, "start")
    for prec in (32, 64):
        t0 = time.time()
        out[f"misfit{prec}"] = oracle.cufd(lam0, mu0, den0, stf, 0, ids, para, precision=prec)[0]
        if not gradient:
            continue
        r = oracle.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=prec)
        for k, name in ((1, "lam"), (2, "mu"), (3, "den")):
            out[f"{name}{prec}"] = sketch(r[k])
            out[f"{name}{prec}_norm"] = np.linalg.norm(r[k])
        out[f"stf{prec}"] = np.asarray(r[4][: len(ids)], dtype=np.float32)
        log(case.name, "gradient", prec, f"{time.time() - t0:.1f}s")
    return out, syn32


def run_c3():
    c = models.case3(nSteps=C3_NT, nshots=64)
    out, _ = grad_case(c, C3_IDS, tempfile.mkdtemp())
    np.savez(os.path.join(HERE, "sketch_c3_grad.npz"), **out)


def fd_directions(c):
    m = (c.nPml + 12, c.nPml + c.nPad)  # below the source/receiver rows, above the pad
    return [[smooth_direction((c.NZ, c.NX), 1000 + 10 * k + q, 8.0, (m, c.nPml))
             for q in range(3)] for k in range(FD_DIRS)]


def run_fd():
    """J(m (1 +- h r)) on the 500x1700 grid for FD_DIRS smooth directions r
    (lambda, mu, rho each perturbed), and the gradient's directional
    derivative sum(g * m * r), in FP64 (and the FP32 misfits)."""
    c = models.case3(nSteps=C3_NT, nshots=64)
    ws = tempfile.mkdtemp()
    para, stf = models.write_case(c, ws)
    lam, mu, den = models.moduli(c, "true")
    oracle.cufd(lam, mu, den, stf, 2, C3_IDS, para, precision=32)
    m0 = models.moduli(c, "start")
    dirs = fd_directions(c)
    out = {"h": FD_H, "ids": np.array(C3_IDS)}
    for prec in (64, 32):
        r = oracle.cufd(*m0, stf, 1, C3_IDS, para, precision=prec)
        g = r[1:4]
        out[f"gdm{prec}"] = np.array([sum(np.sum(gq * mq * dq) for gq, mq, dq in zip(g, m0, d))
                                      for d in dirs])
        Jp, Jm = [], []
        for k, d in enumerate(dirs):
            t0 = time.time()
            mp = [mq * (1.0 + FD_H * dq) for mq, dq in zip(m0, d)]
            mm = [mq * (1.0 - FD_H * dq) for mq, dq in zip(m0, d)]
            Jp.append(oracle.cufd(*mp, stf, 0, C3_IDS, para, precision=prec)[0])
            Jm.append(oracle.cufd(*mm, stf, 0, C3_IDS, para, precision=prec)[0])
            log("fd", prec, k, Jp[-1], Jm[-1], f"{time.time() - t0:.1f}s")
        out[f"Jp{prec}"], out[f"Jm{prec}"] = np.array(Jp), np.array(Jm)
        out[f"fd{prec}"] = (out[f"Jp{prec}"] - out[f"Jm{prec}"]) / (2 * FD_H)
        log("fd", prec, "fd", out[f"fd{prec}"], "gdm", out[f"gdm{prec}"])
    np.savez(os.path.join(HERE, "sketch_c3_fd.npz"), **out)


def run_c4():
    for nz, nx in C4_GRIDS:
        ids = c4_ids(nz, nx)
        syn = {}
        for monitor in (False, True):
            c = models.case4(nz, nx, monitor=monitor, nSteps=C4_NT, nshots=16)
            out, syn[monitor] = grad_case(c, ids, tempfile.mkdtemp())
            np.savez(os.path.join(HERE, f"sketch_{c.name}.npz"), **out)
        d = float(np.linalg.norm(syn[True] - syn[False]) / np.linalg.norm(syn[False]))
        log(f"c4 {nz}x{nx} monitor vs baseline data rel {d:.3e}")
        np.savez(os.path.join(HERE, f"sketch_c4_{nz}x{nx}_timelapse.npz"), monitor_rel=d)


def run_c2():
    """Water layer (Vs = 0 in the top 10 rows): seismograms and misfit only,
    since the reference's adjoint grows without bound over a long record with
    such a layer; elastic top: gradients too."""
    for water in (True, False):
        c = models.case2(water=water)
        out, _ = grad_case(c, C2_IDS, tempfile.mkdtemp(), gradient=not water)
        np.savez(os.path.join(HERE, f"sketch_c2_{'water' if water else 'elastic'}.npz"), **out)


if __name__ == "__main__":
    oracle.build()
    which = sys.argv[1:] or ["c3", "fd", "c4", "c2"]
    if "c3" in which:
        run_c3()
    if "fd" in which:
        run_fd()
    if "c4" in which:
        run_c4()
    if "c2" in which:
        run_c2()
