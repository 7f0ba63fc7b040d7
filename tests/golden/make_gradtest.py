"""The reference's own gradient test, restated: deps/CustomOps/FWI/gradtest.jl.

400x400 grid (20 m, nPml 32, nPad 0), nSteps 2001, dt 2.5 ms, f0 4.5 Hz; one
source at interior index (168, 168), a 13x13 receiver mesh at 5:10:331
(gradtest.jl:16-39); homogeneous Vp 3000, Vs 3000/sqrt(3), rho 2000 for the
observed data; the misfit as a function of mu evaluated at the mu of
Vp 3200, Vs 3200/sqrt(3) (gradtest.jl:112-113); direction v =
(1 + 0.1 rand) * 1e3 on [nPml+4, n-nPml-5) (0-based), zero elsewhere
("Don't perturb in the CPML region", gradtest.jl:122-126; numpy's default_rng
seed 233 here); Taylor test at gamma = 1e-1, 1e-2, 1e-3 (gradtest.jl:137-151).

    python tests/golden/make_gradtest.py

writes tests/golden/gradtest.npz: the oracle's FP32 (reference arithmetic)
and FP64 misfit y0, <v, grad>, s(gamma) = f(m + gamma v) - f(m) and
w(gamma) = s(gamma) - gamma <v, grad>.
"""
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "fwiflow.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

from fwiflow.utils import paraGen, sourceGene, surveyGen  # noqa: E402
from oracle import oracle  # noqa: E402

N, D, NT, DT, F0, NPML = 400, 20.0, 2001, 0.0025, 4.5, 32
GAMMAS = np.array([1e-1, 1e-2, 1e-3])


def setup(ws):
    """Parameter and survey files, moduli, stf, direction (shared with the
    GPU test)."""
    para = os.path.join(ws, "para_file.json")
    survey = os.path.join(ws, "survey_file.json")
    paraGen(N, N, D, D, NT, DT, F0, NPML, 0, para, survey, os.path.join(ws, "Data"))
    r = np.arange(5, N - 2 * NPML - 5, 10)
    xr, zr = np.meshgrid(r, r)
    surveyGen(np.array([200 - NPML]), np.array([200 - NPML]), zr.ravel(), xr.ravel(), survey)

    def moduli(cp):
        cs, den = cp / np.sqrt(3.0), np.full((N, N), 2000.0)
        return (cp ** 2 - 2.0 * cs ** 2) * den / 1e6, cs ** 2 * den / 1e6, den

    lam, mu, den = moduli(np.full((N, N), 3000.0))
    mu2 = moduli(np.full((N, N), 3200.0))[1]
    stf = sourceGene(F0, NT, DT).reshape(1, NT)
    v0 = np.zeros((N, N))
    v0[NPML + 4:N - NPML - 5, NPML + 4:N - NPML - 5] = 1.0
    v = (1.0 + 0.1 * np.random.default_rng(233).random((N, N))) * 1e3 * v0
    return para, lam, mu, den, mu2, stf, v


if __name__ == "__main__":
    oracle.build()
    ws = tempfile.mkdtemp()
    para, lam, mu, den, mu2, stf, v = setup(ws)
    oracle.cufd(lam, mu, den, stf, 2, [0], para, precision=32)
    out = {"gamma": GAMMAS}
    for prec in (32, 64):
        y0 = oracle.cufd(lam, mu2, den, stf, 0, [0], para, precision=prec)[0]
        g = oracle.cufd(lam, mu2, den, stf, 1, [0], para, precision=prec)[2]
        vg = float(np.sum(v * g))
        s = np.array([oracle.cufd(lam, mu2 + gm * v, den, stf, 0, [0], para, precision=prec)[0]
                      - y0 for gm in GAMMAS])
        out[f"y{prec}"], out[f"vg{prec}"] = y0, vg
        out[f"s{prec}"], out[f"w{prec}"] = s, s - GAMMAS * vg
        print(prec, "y0", y0, "vg", vg, "s", s, "w", s - GAMMAS * vg, flush=True)
    np.savez(os.path.join(HERE, "gradtest.npz"), **out)
