"""GPU exact adjoint (fwiflow_plan_set_adjoint(plan, 1), Plan.set_adjoint
("exact"); no reference counterpart) against the oracle's exact adjoint, its
backward scheme 3, which tests/test_oracle.py pins against central
differences of the FP64 misfit (to 1e-7, with waves in the C-PML and with a
zero-shear layer).  Every gradient (lambda, mu, rho, stf) within 1e-3 of the
FP32 oracle and, against FP64, within the tolerance rule of test_gpu.py.
"""
import json
import os

import numpy as np
import pytest

import fwiflow
from fwiflow import models

from test_gpu import _assert, env, rel, say

pytestmark = pytest.mark.gpu

CASES = {
    "ring": {},
    "checkpoints_7": {"checkpoint": 7},
    "zero_shear_layer": {"water": True},
    "groups_of_1": {"settings": {"FWIFLOW_SHOT_GROUP": "1", "FWIFLOW_SHOT_GROUP_BWD": "1"}},
    "irregular_survey": {"irregular": True},
    "graph_replay": {"settings": {"FWIFLOW_GRAPHS": "1"}},
}


def _case(ws, water=False, irregular=False):
    c = models.case_small(nz=40, nx=60, nSteps=300, nshots=3)
    if water:
        for a in (c.cs, c.cs0):
            a[:5] = 0.0
    para, stf = models.write_case(c, ws)
    if irregular:
        # receivers off one row (the separate injection launch), two in one cell
        d = json.loads(open(para).read())
        sv = json.loads(open(d["survey_fname"]).read())
        recs = [(list(range(2, 30)) + [1] * 20, [40] * 28 + list(range(5, 45, 2))),
                ([6] * 11, list(range(10, 54, 4))),
                ([3, 3, 3, 12], [20, 20, 33, 33])]
        for s, (zr, xr) in enumerate(recs):
            sv[f"shot{s}"].update(z_rec=zr, x_rec=xr, nrec=len(zr))
        open(d["survey_fname"], "w").write(json.dumps(sv))
    ids = [0, 1, 2]
    fwiflow.fwi_obs_op(*models.moduli(c, "true"), stf, 0, ids, para)
    return c, para, stf, ids, models.moduli(c, "start")


def _exact_run(para, ids, m0, stf, settings=None, checkpoint=None, mode="exact"):
    with env(**(settings or {})), fwiflow.Plan(para, ids) as P:
        P.set_adjoint(mode)
        P.set_model(*m0)
        P.set_stf(stf)
        P.load_obs()
        if checkpoint is not None:
            P.set_checkpoint(checkpoint)
        P.run(1)
        g, J = P.gradients(), P.misfit()
        if (settings or {}).get("FWIFLOW_GRAPHS") == "1":
            # the captured backward replayed, with eager work in between
            for _ in range(2):
                P.run(2)
                P.run(1)
                assert all(np.array_equal(a, b) for a, b in zip(P.gradients(), g))
        return g, J


@pytest.mark.parametrize("case", list(CASES))
def test_exact_adjoint_matches_oracle(oracle_mod, workspace, case):
    kw = CASES[case]
    c, para, stf, ids, m0 = _case(workspace, kw.get("water", False), kw.get("irregular", False))
    say("exact adjoint", case)
    g, J = _exact_run(para, ids, m0, stf, kw.get("settings"), kw.get("checkpoint"))
    o32 = oracle_mod.cufd(*m0, stf, 1, ids, para, precision=32, scheme=3, ckpt_every=50)
    o64 = oracle_mod.cufd(*m0, stf, 1, ids, para, precision=64, scheme=3, ckpt_every=50)
    Jo = oracle_mod.cufd(*m0, stf, 0, ids, para, precision=32)[0]
    out = {"misfit": (abs(J - Jo) / abs(Jo), None, None)}
    for k, name in enumerate(("lam", "mu", "den", "stf")):
        a, b, d = g[k], o32[k + 1], o64[k + 1]
        if name == "stf":
            a, b, d = a[:3], b[:3], d[:3]
        out[name] = (rel(a, b), rel(a, d), rel(b, d))
    say(case, {k: tuple(None if x is None else f"{x:.2e}" for x in v) for k, v in out.items()})
    _assert(out)


def test_exact_and_reference_modes_differ_where_the_reference_is_inexact(workspace):
    """With a zero-shear layer the two adjoints give different gradients, and
    a plan switched back to the reference's adjoint reproduces a fresh
    reference-mode plan bit for bit (the switch leaves no state behind)."""
    c, para, stf, ids, m0 = _case(workspace, water=True)
    ge, _ = _exact_run(para, ids, m0, stf)
    gr, _ = _exact_run(para, ids, m0, stf, mode="reference")
    d = [rel(a, b) for a, b in zip(ge[:3], gr[:3])]
    say("exact vs reference", [f"{x:.2e}" for x in d])
    assert max(d) > 1e-2
    with fwiflow.Plan(para, ids) as P:
        P.set_model(*m0)
        P.set_stf(stf)
        P.load_obs()
        P.set_adjoint("exact")
        P.run(1)
        P.set_adjoint("reference")
        P.run(1)
        g2 = P.gradients()
    assert all(np.array_equal(a, b) for a, b in zip(g2, gr))
    with pytest.raises(ValueError):
        with fwiflow.Plan(para, ids) as P:
            P.set_adjoint("transpose")


def test_exact_adjoint_in_a_job_runs_plans_one_after_another(workspace):
    """A job holding an exact-adjoint plan runs its plans one after another
    (fused() is false) and gives each plan's own results."""
    cs = [models.case_small(nz=40, nx=60, nSteps=200, nshots=2, seed=1),
          models.case_small(nz=50, nx=80, nSteps=200, nshots=2, seed=2)]
    sep, plans = [], []
    try:
        for k, c in enumerate(cs):
            ws = os.path.join(workspace, str(k))
            para, stf = models.write_case(c, ws)
            ids = [0, 1]
            fwiflow.fwi_obs_op(*models.moduli(c, "true"), stf, 0, ids, para)
            m0 = models.moduli(c, "start")
            sep.append(_exact_run(para, ids, m0, stf))
            P = fwiflow.Plan(para, ids)
            plans.append(P)
            P.set_adjoint("exact")
            P.set_model(*m0)
            P.set_stf(stf)
            P.load_obs()
        with fwiflow.Job(plans) as Jb:
            Jb.run(1)
            assert not Jb.fused()
            for P, (g0, J0) in zip(plans, sep):
                assert P.misfit() == J0
                assert all(np.array_equal(a, b) for a, b in zip(P.gradients(), g0))
    finally:
        for P in plans:
            P.close()


def test_exact_adjoint_with_conditioning_matches_oracle(oracle_mod, workspace):
    """Per-trace windows and weights plus the band-pass filter
    (libCUFD.cu:254-309): the exact adjoint injects the conditioned
    residual, as the reference's does; gradients against the oracle's scheme
    3 with the same stages, FP32 and FP64."""
    from test_gpu import _cond_setup
    c, para, stf, cond = _cond_setup(workspace, if_win=True, filter_para=[2.0, 5.0, 25.0, 40.0])
    ids = list(range(len(c.src_x)))
    g = oracle_mod.geom(c.NZ, c.NX, c.nPml, c.nPad, c.nSteps, c.dz, c.dx, c.dt, c.f0)
    src = [(z + c.nPml, x + c.nPml) for z, x in zip(c.src_z, c.src_x)]
    recs = [(c.rec_z + c.nPml, c.rec_x + c.nPml)] * len(src)
    obs = oracle_mod.run(g, 2, *models.moduli(c, "true"), stf, src, recs, precision=32)["syn"]
    m0 = models.moduli(c, "start")
    o32 = oracle_mod.run(g, 1, *m0, stf, src, recs, obs, precision=32, cond=cond, scheme=3,
                         ckpt_every=50)
    o64 = oracle_mod.run(g, 1, *m0, stf, src, recs, obs, precision=64, cond=cond, scheme=3,
                         ckpt_every=50)
    with fwiflow.Plan(para, ids) as P:
        P.set_adjoint("exact")
        P.set_model(*m0)
        P.set_stf(stf)
        P.set_obs(obs)
        P.run(1)
        gL, gM, gD, gS = P.gradients()
    out = {}
    n = len(ids)
    for name, a, key in (("lam", gL, "grad_lambda"), ("mu", gM, "grad_mu"), ("den", gD, "grad_den"),
                         ("stf", gS[:n], "grad_stf")):
        b, d = np.asarray(o32[key]), np.asarray(o64[key])
        if name == "stf":  # rows 0..n-1 hold the group's shots
            b, d = b[:n], d[:n]
        out[name] = (rel(a, b), rel(a, d), rel(b, d))
    say("exact + conditioning", {k: tuple(f"{x:.2e}" for x in v) for k, v in out.items()})
    _assert(out)


def test_exact_adjoint_through_cufd_and_noalloc(workspace):
    """FWIFLOW_EXACT_ADJOINT=1 selects the exact adjoint for the stateless
    entry points too: fwiflow_cufd and fwiflow_cufd_noalloc give the plan's
    exact gradients (to the summation order of the shots of a launch)."""
    c, para, stf, ids, m0 = _case(workspace)
    g, J = _exact_run(para, ids, m0, stf)
    for noalloc in (False, True):
        with env(FWIFLOW_EXACT_ADJOINT="1"):
            _, gL, gM, gD, gS = fwiflow.ops.cufd(*m0, stf, 1, 0, ids, para, noalloc=noalloc)
        d = [rel(a, b) for a, b in zip((gL, gM, gD, gS), g)]
        say("cufd exact", "noalloc" if noalloc else "", [f"{x:.1e}" for x in d])
        assert max(d) <= 1e-6, d
    _, gL, _, _, _ = fwiflow.ops.cufd(*m0, stf, 1, 0, ids, para)
    assert rel(gL, g[0]) > 1e-3  # without the variable: the reference's adjoint
