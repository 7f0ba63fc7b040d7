"""Config writers and helpers of the reference (src/Utils.jl, src/FWI.jl).

* ``paraGen`` / ``surveyGen`` write the one-line JSON files the operator reads
  (Utils.jl:21-109).
* ``sourceGene`` integrated Ricker wavelet (Utils.jl:116-132).
* ``velocity_to_moduli`` (Utils.jl:221-227), ``padding`` (FWI.jl:193-205,
  TF "SYMMETRIC" pad == numpy "symmetric").
"""
from __future__ import annotations

import json
import os
from collections import OrderedDict

import numpy as np


def paraGen(nz, nx, dz, dx, nSteps, dt, f0, nPml, nPad, para_fname, survey_fname,
            data_dir_name, if_win=False, filter_para=None, if_src_update=False,
            scratch_dir_name=""):
    """Utils.jl:21-63.  Creates data_dir_name (and scratch_dir_name) if missing."""
    para = OrderedDict()
    para["nz"] = int(nz)
    para["nx"] = int(nx)
    para["dz"] = float(dz)
    para["dx"] = float(dx)
    para["nSteps"] = int(nSteps)
    para["dt"] = float(dt)
    para["f0"] = float(f0)
    para["nPoints_pml"] = int(nPml)
    para["nPad"] = int(nPad)
    if if_win:
        para["if_win"] = True
    if filter_para is not None:
        para["filter"] = [float(v) for v in filter_para]
    if if_src_update:
        para["if_src_update"] = True
    para["survey_fname"] = survey_fname
    para["data_dir_name"] = data_dir_name
    os.makedirs(data_dir_name, exist_ok=True)
    if scratch_dir_name != "":
        para["scratch_dir_name"] = scratch_dir_name
        os.makedirs(scratch_dir_name, exist_ok=True)
    with open(para_fname, "w") as f:
        f.write(json.dumps(para))


def surveyGen(z_src, x_src, z_rec, x_rec, survey_fname, Windows=None, Weights=None):
    """Utils.jl:80-109: every shot gets the same receiver line."""
    z_src, x_src = np.atleast_1d(z_src), np.atleast_1d(x_src)
    z_rec, x_rec = np.atleast_1d(z_rec), np.atleast_1d(x_rec)
    survey = OrderedDict()
    survey["nShots"] = int(len(x_src))
    for i in range(len(x_src)):
        shot = OrderedDict()
        shot["z_src"] = int(z_src[i])
        shot["x_src"] = int(x_src[i])
        shot["nrec"] = int(len(x_rec))
        shot["z_rec"] = [int(v) for v in z_rec]
        shot["x_rec"] = [int(v) for v in x_rec]
        if Windows is not None:
            shot["win_start"] = list(Windows[f"shot{i}"]["start"])
            shot["win_end"] = list(Windows[f"shot{i}"]["end"])
        if Weights is not None:
            shot["weights"] = list(Weights[f"shot{i}"]["weights"])
        survey[f"shot{i}"] = shot
    with open(survey_fname, "w") as f:
        f.write(json.dumps(survey))


def sourceGene(f, nStep, delta_t):
    """Utils.jl:116-132: Ricker (1-2e tau^2) exp(-e tau^2), e = pi^2 f^2,
    tau = dt*(it-1) - 1.2/f, cumulatively summed and scaled by dt.  Shape (1, nStep)."""
    e = np.pi * np.pi * f * f
    t_delay = 1.2 / f
    it = np.arange(1, nStep + 1)
    tau = delta_t * (it - 1) - t_delay
    src = (1 - 2 * e * tau ** 2) * np.exp(-e * tau ** 2)
    # Julia's loop: source[it] = source[it] + source[it-1] (left-to-right prefix sum)
    src = np.cumsum(src)
    return (src * delta_t).reshape(1, nStep)


def velocity_to_moduli(cp, cs, den):
    """Utils.jl:221-227: lambda = (cp^2 - 2 cs^2) rho / 1e6, mu = cs^2 rho / 1e6 (MPa)."""
    cp, cs, den = (np.asarray(a, dtype=np.float64) for a in (cp, cs, den))
    lam = (cp * cp - 2.0 * cs * cs) * den / 1e6
    mu = cs * cs * den / 1e6
    return lam, mu


def moduli_to_velocity_grad(cp, cs, den, g_lam, g_mu, g_den):
    """Chain rule of velocity_to_moduli: (dJ/dcp, dJ/dcs, dJ/dden) from
    (dJ/dlambda, dJ/dmu, dJ/dden_direct)."""
    cp, cs, den = (np.asarray(a, dtype=np.float64) for a in (cp, cs, den))
    g_cp = g_lam * 2.0 * cp * den / 1e6
    g_cs = (g_lam * (-4.0) * cs + g_mu * 2.0 * cs) * den / 1e6
    g_rho = g_den + g_lam * (cp * cp - 2.0 * cs * cs) / 1e6 + g_mu * cs * cs / 1e6
    return g_cp, g_cs, g_rho


def npad(nz, nPml):
    """FWI.jl:12: nPad = 32 - mod(nz + 2 nPml, 32) (32 when already aligned)."""
    return 32 - ((nz + 2 * nPml) % 32)


def padding(a, nPml, nPad):
    """FWI.jl:202: symmetric pad [nPml, nPml+nPad] rows, [nPml, nPml] columns."""
    return np.pad(np.asarray(a, dtype=np.float64), ((nPml, nPml + nPad), (nPml, nPml)),
                  mode="symmetric")


def padding_adjoint(g, nz, nx, nPml, nPad):
    """Adjoint of ``padding``: folds the padded gradient back onto (nz, nx)."""
    idx = np.arange(nz * nx).reshape(nz, nx)
    pidx = np.pad(idx, ((nPml, nPml + nPad), (nPml, nPml)), mode="symmetric")
    out = np.zeros(nz * nx)
    np.add.at(out, pidx.ravel(), np.asarray(g, dtype=np.float64).ravel())
    return out.reshape(nz, nx)
