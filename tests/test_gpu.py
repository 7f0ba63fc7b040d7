"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Tolerances (stated by the operator's contract): seismograms and misfits
within 1e-4 relative L2 of the oracle run in the reference's FP32
arithmetic, every gradient (lambda, mu, rho and the source-time function)
within 1e-3.  Against the FP64 oracle, each quantity is held to the contract
tolerance or, where FP32 itself is farther from FP64 than that, to the FP32
oracle's own distance from FP64 (its "spread") plus twice the GPU's distance
from the FP32 oracle.
"""
import contextlib
import os
import sys

import numpy as np
import pytest

import fwiflow
from fwiflow import models
from fwiflow.models import Case

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-4
GRAD_TOL = 1e-3
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@contextlib.contextmanager
def env(**kw):
    """Engine settings are read when a plan is created: set them around it."""
    old = {k: os.environ.get(k) for k in kw}
    os.environ.update({k: str(v) for k, v in kw.items()})
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _obs_files(para_dir, ids, nrec, nt):
    return [np.fromfile(os.path.join(para_dir, "Data", f"Shot{i}.bin"),
                        dtype=np.float32).reshape(nrec, nt) for i in ids]


def _parity(oracle_mod, case, ws, ids=None, check_fp64=True, gradient=True):
    """Observed data from the true model on the GPU (their Shot files are the
    observed data of both sides), then misfit and gradients at the starting
    model on the GPU and the oracle.  Returns {quantity: (rel to FP32 oracle,
    rel to FP64 oracle or None, FP32 oracle's spread from FP64 or None)}."""
    para, stf = models.write_case(case, ws)
    ids = list(range(len(case.src_x))) if ids is None else ids
    lam, mu, den = models.moduli(case, "true")
    say("case", case.name, "grid", case.NZ, case.NX, "nt", case.nSteps, "shots", ids)
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, ids, para)
    nrec = len(case.rec_x)
    gpu_obs = np.concatenate(_obs_files(ws, ids, nrec, case.nSteps))
    o32 = np.concatenate(oracle_mod.cufd(lam, mu, den, stf, 2, ids, para, precision=32)[5]["syn"])
    out = {}

    def put(name, g, a32, a64):
        e32 = rel(g, a32)
        if a64 is None:
            out[name] = (e32, None, None)
        else:
            out[name] = (e32, rel(g, a64), rel(a32, a64))

    o64 = None
    if check_fp64:
        o64 = np.concatenate(
            oracle_mod.cufd(lam, mu, den, stf, 2, ids, para, precision=64)[5]["syn"])
    put("seis", gpu_obs, o32, o64)
    lam0, mu0, den0 = models.moduli(case, "start")
    J = fwiflow.fwi_op(lam0, mu0, den0, stf, 0, ids, para)
    Jo = oracle_mod.cufd(lam0, mu0, den0, stf, 0, ids, para, precision=32)[0]
    J64 = (oracle_mod.cufd(lam0, mu0, den0, stf, 0, ids, para, precision=64)[0]
           if check_fp64 else None)
    put("misfit", [J], [Jo], None if J64 is None else [J64])
    if gradient:
        G = fwiflow.fwi_op_grad(lam0, mu0, den0, stf, 0, ids, para)
        r = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=32)
        r64 = (oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=64)
               if check_fp64 else None)
        n = len(ids)
        for k, name in enumerate(("lam", "mu", "den", "stf")):
            a, b = G[k], r[k + 1]
            c = None if r64 is None else r64[k + 1]
            if name == "stf":  # rows 0..n-1 hold the group's shots
                a, b, c = a[:n], b[:n], None if c is None else c[:n]
            put(name, a, b, c)
    say(case.name, {k: tuple(None if x is None else f"{x:.2e}" for x in v)
                    for k, v in out.items()})
    return out


def _assert(out):
    for k, (e32, e64, spread) in out.items():
        tol = SEIS_TOL if k in ("seis", "misfit") else GRAD_TOL
        assert e32 <= tol, (k, "FP32", e32, tol)
        if e64 is not None:
            tol64 = max(tol, spread + 2.0 * e32 + 1e-9)
            assert e64 <= tol64, (k, "FP64", e64, tol64)


def test_small_case_parity(oracle_mod, workspace):
    _assert(_parity(oracle_mod, models.case_small(nSteps=300), workspace))


def test_config1_parity(oracle_mod, workspace):
    """100x200 + 32-cell C-PML, nt 1000, dt 2 ms, 1 shot, 200 receivers, 10 Hz."""
    _assert(_parity(oracle_mod, models.case1(), workspace))


@pytest.mark.timeout(900)
def test_docs_marmousi_parity(oracle_mod, workspace):
    """The reference's tutorial setup (docs/src/tutorials/fwi.md:12-43): the
    Marmousi Vp of docs/data (interior of the 224x448 file, re-padded),
    acoustic (cs = 0), rho = 2500, FWIExample acquisition written as the
    survey file stores it (ind_src_z = 2, ind_src_x = 4:8:384, receivers
    3:381; FWI.jl:85-95), its 4.5 Hz source file; 3 of its 48 shots and the
    first 800 of 2000 steps, FP32 and FP64 oracles."""
    nz, nx, nPml = 134, 384, 32
    NZ, NX = 224, 448
    cp = np.fromfile(os.path.join(GOLDEN, "Model_Cp_true.bin"), dtype=np.float32)
    cp = cp.astype(np.float64).reshape(NX, NZ).T  # z fastest on disk
    cp0 = np.fromfile(os.path.join(GOLDEN, "Model_Cp_init_1D.bin"), dtype=np.float32)
    cp0 = cp0.astype(np.float64).reshape(NX, NZ).T
    stf = np.fromfile(os.path.join(GOLDEN, "sourceF_4p5_2_high.bin"), dtype=np.float32)
    nt = 800
    inner = (slice(nPml, nPml + nz), slice(nPml, nPml + nx))
    src_x = np.arange(4, 385, 8)
    case = Case("marmousi", nz, nx, 24.0, 24.0, nt, 0.0025, 4.5, nPml, cp[inner],
                np.zeros((nz, nx)), np.full((nz, nx), 2500.0), cp0[inner], np.zeros((nz, nx)),
                np.full((nz, nx), 2500.0), 2 * np.ones(len(src_x), int), src_x,
                2 * np.ones(379, int), np.arange(3, 382),
                stf[:nt].astype(np.float64).reshape(1, nt))
    assert case.NZ == NZ and case.NX == NX
    _assert(_parity(oracle_mod, case, workspace, ids=[0, 9, 40]))


def test_batched_shots_equal_separate_runs(workspace):
    c = models.case_small(nSteps=250, nshots=3)
    para, stf = models.write_case(c, workspace)
    lam, mu, den = models.moduli(c, "true")
    fwiflow.fwi_obs_op(lam * 1.01, mu, den, stf, 0, [0, 1, 2], para)
    with fwiflow.Plan(para, [0, 1, 2]) as P:
        P.set_model(lam, mu, den)
        P.set_stf(stf)
        P.load_obs()
        P.run(2)
        batched = P.data()
        P.run(1)
        gb = P.gradients()
    sep, gs = [], None
    for i in range(3):
        with fwiflow.Plan(para, [i]) as P:
            P.set_model(lam, mu, den)
            P.set_stf(stf)
            P.load_obs()
            P.run(2)
            sep.append(P.data())
            P.run(1)
            g = P.gradients()
            gs = list(g[:3]) if gs is None else [a + b for a, b in zip(gs, g[:3])]
    assert np.array_equal(batched, np.concatenate(sep))  # shots are independent
    for a, b in zip(gb[:3], gs):
        assert rel(a, b) < 1e-5


@pytest.mark.parametrize("graphs", ["0", "1"], ids=["eager", "graphs"])
def test_plan_rerun_is_deterministic(workspace, graphs):
    """Repeated gradient runs of one plan, with forward-only and data runs
    between them, give the same bits; with captured graphs every replay
    (the first run captures) equals the eager plan's result."""
    c = models.case_small(nSteps=200)
    para, stf = models.write_case(c, workspace)
    lam, mu, den = models.moduli(c, "true")
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, [0, 1], para)
    lam0, mu0, den0 = models.moduli(c, "start")
    outs = []
    with env(FWIFLOW_GRAPHS=graphs), fwiflow.Plan(para, [0, 1]) as P:
        P.set_model(lam0, mu0, den0)
        P.set_stf(stf)
        P.load_obs()
        for _ in range(3):
            P.run(1)
            outs.append((P.misfit(), P.gradients()))
            P.run(0)
            J = P.misfit()
            P.run(2)
    for Jk, gk in outs[1:]:
        assert Jk == outs[0][0]
        for x, y in zip(gk, outs[0][1]):
            assert np.array_equal(x, y)
    assert J == fwiflow.fwi_op(lam0, mu0, den0, stf, 0, [0, 1], para)
    g_eager = fwiflow.fwi_op_grad(lam0, mu0, den0, stf, 0, [0, 1], para)
    for x, y in zip(outs[0][1], g_eager):
        assert np.array_equal(x, y)


def test_misfit_zero_at_truth_and_zfast_layout(workspace):
    c = models.case_small(nSteps=200)
    para, stf = models.write_case(c, workspace)
    lam, mu, den = models.moduli(c, "true")
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, [0, 1], para)
    assert fwiflow.fwi_op(lam, mu, den, stf, 0, [0, 1], para) == 0.0
    with fwiflow.Plan(para, [0, 1]) as P:  # Julia-native column-major input
        P.set_model(lam.ravel(order="F"), mu.ravel(order="F"), den.ravel(order="F"),
                    zfast=True)
        P.set_stf(stf)
        P.load_obs()
        P.run(0)
        assert P.misfit() == 0.0


def test_full_size_grid_properties(oracle_mod, workspace):
    """The 500x1700 (+40 C-PML) grid at full size, 2 shots, first 300 steps:
    seismograms against the oracle, linearity in the source, misfit 0 at truth."""
    c = models.case3(nSteps=300, nshots=64)
    para, stf = models.write_case(c, workspace)
    lam, mu, den = models.moduli(c, "true")
    ids = [0, 40]
    say("case c3-grid", c.NZ, c.NX, "nt", c.nSteps)
    with fwiflow.Plan(para, ids) as P:
        P.set_model(lam, mu, den)
        P.set_stf(stf)
        P.run(2)
        d1 = P.data()
        P.set_stf(-3.0 * stf)
        P.run(2)
        d3 = P.data()
    assert rel(d3, -3.0 * d1) < 1e-6
    o = oracle_mod.cufd(lam, mu, den, stf, 2, ids, para, precision=32)[5]
    ref = np.concatenate([s.ravel() for s in o["syn"]])
    assert rel(d1, ref) < SEIS_TOL
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, ids, para)
    assert fwiflow.fwi_op(lam, mu, den, stf, 0, ids, para) == 0.0


def test_velocity_api_and_autograd(workspace):
    """compute_misfit_and_gradient (FWI.jl's chain: pad, mask, moduli) and the
    torch autograd op: backward = gradient x dy."""
    torch = pytest.importorskip("torch")
    c = models.case_small(nSteps=200)
    fwi = fwiflow.make_fwi(c.nz, c.nx, c.dz, c.dx, c.nSteps, c.dt, c.src_z, c.src_x, c.rec_z,
                           c.rec_x, nPml=c.nPml, f0=c.f0, WORKSPACE=workspace)
    obs = fwiflow.compute_observation(fwi, c.cp, c.cs, c.rho, c.stf)
    assert obs.shape == (2, c.nSteps, c.nx)
    J, gcp, gcs, grho = fwiflow.compute_misfit_and_gradient(
        fwi, c.cp0, c.cs0, c.rho0, c.stf, cp_ref=c.cp0, cs_ref=c.cs0, rho_ref=c.rho0)
    assert gcp.shape == (c.nz, c.nx) and J > 0
    assert J == fwiflow.compute_misfit(fwi, c.cp0, c.cs0, c.rho0, c.stf, cp_ref=c.cp0,
                                       cs_ref=c.cs0, rho_ref=c.rho0)
    # top 10 rows of the interior are masked (FWI.jl:49)
    assert np.all(gcp[:10] == 0) and np.any(gcp[10:] != 0)
    lam0, mu0, den0 = (torch.tensor(a, requires_grad=True) for a in models.moduli(c, "start"))
    stf = torch.tensor(np.repeat(c.stf, 2, axis=0), requires_grad=True)
    loss = 2.5 * fwiflow.ops.FwiOp.apply(lam0, mu0, den0, stf, 0, [0, 1], fwi.para_path)
    loss.backward()
    gL, gM, gD, gS = fwiflow.fwi_op_grad(lam0.detach().numpy(), mu0.detach().numpy(),
                                         den0.detach().numpy(), stf.detach().numpy(), 0, [0, 1],
                                         fwi.para_path)
    assert np.allclose(lam0.grad.numpy(), 2.5 * gL) and np.allclose(stf.grad.numpy(), 2.5 * gS)


def test_illumination_matches_oracle(oracle_mod, workspace):
    """Forward illumination (sum over shots and steps of vz^2 + vx^2) from the
    split forward kernels against the FP32 and FP64 oracles, with shot groups
    of 1 and of all shots, and unchanged data when it is enabled."""
    c = models.case_small(nSteps=200)
    para, stf = models.write_case(c, workspace)
    lam, mu, den = models.moduli(c, "true")
    ids = [0, 1]
    o32 = oracle_mod.cufd(lam, mu, den, stf, 2, ids, para, precision=32, illumination=True)[5]
    o64 = oracle_mod.cufd(lam, mu, den, stf, 2, ids, para, precision=64, illumination=True)[5]
    for group in ("1", "0"):
        with env(FWIFLOW_SHOT_GROUP=group), fwiflow.Plan(para, ids) as P:
            P.set_model(lam, mu, den)
            P.set_stf(stf)
            P.run(2)
            d0 = P.data()
            P.set_illumination(True)
            P.run(2)
            il = P.illumination()
            assert np.array_equal(P.data(), d0)
        e32, e64 = rel(il, o32["illum"]), rel(il, o64["illum"])
        say("illumination group", group, "rel32", e32, "rel64", e64)
        assert e32 < SEIS_TOL and e64 < SEIS_TOL
        assert np.all(il[o64["illum"] == 0.0] == 0.0)


def test_plain_inversion_reduces_misfit(workspace):
    """The L-BFGS-B driver on the device-resident plan: its one-run objective
    equals the two-call compute_misfit / compute_misfit_and_gradient path, a
    few iterations cut the misfit, and every iteration logs a (gradient,
    illumination, update) pair."""
    c = models.case_small(nz=40, nx=60, nSteps=400, nshots=3)
    fwi = fwiflow.make_fwi(c.nz, c.nx, c.dz, c.dx, c.nSteps, c.dt, c.src_z, c.src_x,
                           c.rec_z, c.rec_x, nPml=c.nPml, f0=c.f0, WORKSPACE=workspace)
    fwiflow.compute_observation(fwi, c.cp, c.cs, c.rho, c.stf)
    cp0 = fwiflow.padding(c.cp0, fwi.nPml, fwi.nPad)
    obj = fwiflow.PlanObjective(fwi, c.cs, c.rho, c.stf, cp_ref=c.cp0)
    try:
        J0, g0, ex = obj(cp0)
        Jt = fwiflow.compute_misfit(fwi, cp0, c.cs, c.rho, c.stf, cp_ref=c.cp0, cs_ref=c.cs,
                                    rho_ref=c.rho)
        _, gt, _, _ = fwiflow.compute_misfit_and_gradient(fwi, cp0, c.cs, c.rho, c.stf,
                                                         cp_ref=c.cp0, cs_ref=c.cs,
                                                         rho_ref=c.rho)
        say("plan objective", J0, "two-call", Jt)
        assert J0 == Jt
        assert rel(g0, gt.ravel()) < 1e-6
        assert np.all(ex["illumination"] >= 0.0) and ex["illumination"].max() > 0.0
        log = os.path.join(workspace, "inv")
        r = fwiflow.lbfgs(obj, cp0, maxiter=4, log_dir=log)
    finally:
        obj.close()
    say("inversion", r["iterations"], "iterations", r["evaluations"], "evaluations",
        "misfit", J0, "->", r["misfit"], "in", round(r["seconds"], 2), "s")
    assert r["misfit"] < 0.7 * J0
    x = r["x"]
    assert x.min() >= 1500.0 and x.max() <= 5500.0
    pairs = sorted(os.listdir(os.path.join(log, "pairs")))
    assert len(pairs) == r["iterations"] >= 1
    p = np.load(os.path.join(log, "pairs", pairs[0]))
    assert p["gradient"].shape == p["illumination"].shape == p["update"].shape == cp0.shape
    np.testing.assert_allclose(p["gradient"], g0.reshape(cp0.shape).astype(np.float32),
                               rtol=1e-6, atol=1e-6 * np.abs(g0).max())


def _grad_setup(ws, nz=40, nx=61, nSteps=240, nshots=3, scale=1.02):
    c = models.case_small(nz=nz, nx=nx, nSteps=nSteps, nshots=nshots)
    para, stf = models.write_case(c, ws)
    lam, mu, den = models.moduli(c, "true")
    ids = list(range(nshots))
    fwiflow.fwi_obs_op(lam * scale, mu, den, stf, 0, ids, para)
    return c, para, stf, ids, models.moduli(c, "start")


def _run_plan(para, ids, lam, mu, den, stf, settings=None, checkpoint=None, budget=None):
    with env(**(settings or {})), fwiflow.Plan(para, ids) as P:
        P.set_model(lam, mu, den)
        P.set_stf(stf)
        P.load_obs()
        if checkpoint is not None:
            P.set_checkpoint(checkpoint)
        if budget is not None:
            P.set_memory_budget(budget)
        P.run(2)
        d = P.data()
        P.run(1)
        g = P.gradients()
        plan = P.memory_plan()
        P.run(0)
        J = P.misfit()
    return d, g, J, plan


@pytest.mark.parametrize("settings", [{}, {"FWIFLOW_SHOT_GROUP": "2", "FWIFLOW_SHOT_GROUP_BWD": "1"}],
                         ids=["default_groups", "groups_2_fwd_1_bwd"])
def test_interior_checkpoints_match_oracle(oracle_mod, workspace, settings):
    """Gradient runs that keep the interior states every k steps (k = 7, a
    ragged last interval) and recompute each k-step segment for the reverse
    sweep: gradients against the oracle's scheme 2 (libCUFD.cu:401-404
    interior resets from stored states) in FP32 and FP64, also with the
    checkpoints written by forward groups of 2 and read by backward groups
    of 1."""
    c, para, stf, ids, (lam0, mu0, den0) = _grad_setup(workspace, nz=40, nx=60, nSteps=300,
                                                       scale=1.0)
    lam, mu, den = models.moduli(c, "true")
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, ids, para)
    _, g, _, plan = _run_plan(para, ids, lam0, mu0, den0, stf, settings, checkpoint=7)
    assert plan[0] == 7 and plan[3] > 0
    o32 = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=32, scheme=2,
                          ckpt_every=7)
    o64 = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=64, scheme=2,
                          ckpt_every=7)
    out = {}
    for k, name in enumerate(("lam", "mu", "den", "stf")):
        a, b, d = g[k], o32[k + 1], o64[k + 1]
        if name == "stf":
            a, b, d = a[:3], b[:3], d[:3]
        out[name] = (rel(a, b), rel(a, d), rel(b, d))
    say("checkpoint k=7", settings, {k: tuple(f"{x:.2e}" for x in v) for k, v in out.items()})
    _assert(out)


def test_memory_budget_picks_sweeps_and_checkpoints(workspace):
    """A memory budget too small for the ring of every step of all shots
    splits the gradient run into sweeps of fewer shots (same numbers: each
    shot's data and gradient contributions are computed the same way), and a
    budget below one backward group's full ring switches to checkpoints at
    an interval near sqrt(steps x checkpoint size / ring step size), whose
    gradients equal the explicit-interval run's bit for bit."""
    with env(FWIFLOW_SHOT_GROUP="1"):
        c, para, stf, ids, m0 = _grad_setup(workspace, nSteps=240, nshots=4)
        d0, g0, J0, p0 = _run_plan(para, ids, *m0, stf)
        assert p0[0] == 0 and p0[1] == 4 and p0[3] == 0
        ring_shot = p0[2] / 4
        # two shots' rings fit: sweeps of 2 shots
        d1, g1, J1, p1 = _run_plan(para, ids, *m0, stf, budget=2.5 * ring_shot)
        say("plans", p0, p1)
        assert p1[0] == 0 and p1[1] == 2 and p1[2] < p0[2]
        assert np.array_equal(d1, d0)
        assert J1 == pytest.approx(J0, rel=1e-12)
        for a, b in zip(g1, g0):
            assert rel(a, b) < 1e-6
        # not even one shot's ring fits: checkpoints
        d2, g2, J2, p2 = _run_plan(para, ids, *m0, stf, budget=0.3 * ring_shot)
        say("checkpoint plan", p2)
        k = p2[0]
        assert 1 <= k < c.nSteps - 1 and p2[3] > 0 and p2[2] + p2[3] < p0[2]
        _, g3, _, p3 = _run_plan(para, ids, *m0, stf, checkpoint=k, budget=0.3 * ring_shot)
        assert p3[:2] == p2[:2]
        for a, b in zip(g2, g3):
            assert np.array_equal(a, b)
        for a, b in zip(g2, g0):
            assert rel(a, b) < 1e-3


VARIANTS = {
    "group_all": {"FWIFLOW_SHOT_GROUP": "0"},
    "group2_ragged": {"FWIFLOW_SHOT_GROUP": "2"},
    "group1": {"FWIFLOW_SHOT_GROUP": "1"},
    "unroll1_fwd": {"FWIFLOW_SHOT_UNROLL_FWD": "1"},
    "unroll2_bwd": {"FWIFLOW_SHOT_UNROLL": "2", "FWIFLOW_SHOT_GROUP": "0"},
    "groups_2_fwd_1_bwd": {"FWIFLOW_SHOT_GROUP": "2", "FWIFLOW_SHOT_GROUP_BWD": "1"},
    "tile_x1": {"FWIFLOW_TILE_X": "1"},
    "tile_x2_unbalanced": {"FWIFLOW_TILE_X": "2", "FWIFLOW_XCD_BALANCE": "0"},
    "no_xcd_remap": {"FWIFLOW_XCD_REMAP": "0"},
    "pairs_in_order": {"FWIFLOW_PAIR_MIX": "0", "FWIFLOW_PAIR_MIX_B": "0"},
    "unpaired": {"FWIFLOW_PAIR_A": "0", "FWIFLOW_PAIR_B": "0"},
    "separate_injection": {"FWIFLOW_FOLD_INJECT": "0"},
    "part_table": {"FWIFLOW_PART_TABLE": "1"},
    "graphs": {"FWIFLOW_GRAPHS": "1"},
}


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_launch_variants_match_default(workspace, variant):
    """Every launch arrangement gives the default's numbers: shot groups of
    all shots, ragged groups of 2, single shots, one shot per thread in the
    forward kernels, two shots per thread in the backward kernels,
    backward groups that differ from the forward ones, tiles 1 or 2
    columns wide, the workgroup map without XCD balancing or without XCD
    remapping, the backward stages in order instead of interleaved or as
    four launches per step instead of two pairs, the
    residual injected by a launch of its own, the shot table read from
    device memory instead of kernel arguments, and time loops replayed
    from captured graphs.  Data and gradients within 1e-6 / 1e-5 relative
    (the stencils are per cell and per shot and identical in every variant;
    a launch sums its shots' gradient contributions before adding them to
    the accumulators, so only shot groupings change the order of the sums),
    every variant that keeps the groups bit for bit."""
    c, para, stf, ids, m0 = _grad_setup(workspace)
    d0, g0, J0, _ = _run_plan(para, ids, *m0, stf)
    d1, g1, J1, _ = _run_plan(para, ids, *m0, stf, VARIANTS[variant])
    say(variant, "data", rel(d1, d0), "misfit", abs(J1 - J0) / J0,
        "grads", [rel(a, b) for a, b in zip(g1, g0)])
    assert np.array_equal(d1, d0)
    if variant in ("separate_injection", "part_table", "graphs", "no_xcd_remap",
                   "pairs_in_order", "tile_x1", "tile_x2_unbalanced", "unroll1_fwd",
                   "unpaired"):
        assert J1 == J0 and all(np.array_equal(a, b) for a, b in zip(g1, g0))
    assert abs(J1 - J0) <= 1e-12 * abs(J0)
    for a, b in zip(g1, g0):
        assert rel(a, b) < 1e-5


def test_job_of_two_grids_equals_separate_plans(workspace):
    """A job runs the shots of two plans on different grids (40x60 and 50x80,
    3 and 2 shots, same nSteps) in shared time-step launches: misfits, data
    and gradients equal each plan run alone, bit for bit."""
    cs = [models.case_small(nz=40, nx=60, nSteps=300, nshots=3, seed=1),
          models.case_small(nz=50, nx=80, nSteps=300, nshots=2, seed=2)]
    sep, plans = [], []
    try:
        for k, c in enumerate(cs):
            ws = os.path.join(workspace, str(k))
            para, stf = models.write_case(c, ws)
            lam, mu, den = models.moduli(c, "true")
            ids = list(range(len(c.src_x)))
            fwiflow.fwi_obs_op(lam, mu, den, stf, 0, ids, para)
            m0 = models.moduli(c, "start")
            with fwiflow.Plan(para, ids) as P:
                P.set_model(*m0)
                P.set_stf(stf)
                P.load_obs()
                P.run(1)
                sep.append((P.misfit(), P.gradients(), P.data()))
            P = fwiflow.Plan(para, ids)
            plans.append(P)
            P.set_model(*m0)
            P.set_stf(stf)
            P.load_obs()
        with fwiflow.Job(plans) as J:
            J.run(1)
            assert J.fused()  # both plans' own launches hold all their shots
            for P, (J0, g0, d0) in zip(plans, sep):
                say("job misfit", P.misfit(), "plan", J0)
                assert P.misfit() == J0
                assert np.array_equal(P.data(), d0)
                assert all(np.array_equal(a, b) for a, b in zip(P.gradients(), g0))
    finally:
        for P in plans:
            P.close()


# ------------------------------------------------------------ data conditioning
COND_VARIANTS = {
    "windows": dict(if_win=True),
    "bandpass": dict(filter_para=[2.0, 5.0, 25.0, 40.0]),
    "bandpass_nt_multiple_of_32": dict(filter_para=[2.0, 5.0, 25.0, 40.0], nSteps=320),
    "source_update": dict(if_src_update=True),
    "all_three": dict(if_win=True, filter_para=[2.0, 5.0, 25.0, 40.0], if_src_update=True),
}


def _cond_setup(ws, if_win=False, filter_para=None, if_src_update=False, nSteps=300,
                scratch=False):
    from fwiflow.utils import paraGen, surveyGen
    c = models.case_small(nz=40, nx=60, nSteps=nSteps, nshots=3)
    ns, nrec = len(c.src_x), len(c.rec_x)
    T = c.nSteps * c.dt
    win = {f"shot{s}": {"start": list(0.02 + 0.001 * np.arange(nrec) + 0.01 * s),
                        "end": list(np.full(nrec, 0.85 * T) - 0.003 * s)} for s in range(ns)}
    wts = {f"shot{s}": {"weights": list(1.0 + 0.01 * np.arange(nrec) + 0.1 * s)}
           for s in range(ns)}
    os.makedirs(ws, exist_ok=True)
    para = os.path.join(ws, "para_file.json")
    survey = os.path.join(ws, "survey_file.json")
    paraGen(c.NZ, c.NX, c.dz, c.dx, c.nSteps, c.dt, c.f0, c.nPml, c.nPad, para, survey,
            os.path.join(ws, "Data"), if_win=if_win, filter_para=filter_para,
            if_src_update=if_src_update, scratch_dir_name=os.path.join(ws, "scr") if scratch else "")
    surveyGen(c.src_z, c.src_x, c.rec_z, c.rec_x, survey, Windows=win if if_win else None,
              Weights=wts if if_win else None)
    stf = np.repeat(c.stf.reshape(1, -1), ns, axis=0)
    cond = None
    if if_win or filter_para is not None or if_src_update:
        cond = {"if_win": if_win, "filter": filter_para, "if_src_update": if_src_update,
                "win_start": [np.array(win[f"shot{s}"]["start"]) for s in range(ns)],
                "win_end": [np.array(win[f"shot{s}"]["end"]) for s in range(ns)],
                "weights": [np.array(wts[f"shot{s}"]["weights"]) for s in range(ns)]}
    return c, para, stf, cond


@pytest.mark.parametrize("variant", list(COND_VARIANTS))
def test_conditioning_parity(oracle_mod, workspace, variant):
    """Per-trace windows and weights, band-pass and source update
    (libCUFD.cu:254-309) on the GPU against the oracle: misfit, adjoint source
    (residual after the stages), updated source rows and gradients."""
    kw = COND_VARIANTS[variant]
    c, para, stf, cond = _cond_setup(workspace, **kw)
    ids = list(range(len(c.src_x)))
    say("conditioning", variant, "grid", c.NZ, c.NX, "nt", c.nSteps)
    g = oracle_mod.geom(c.NZ, c.NX, c.nPml, c.nPad, c.nSteps, c.dz, c.dx, c.dt, c.f0)
    src = [(z + c.nPml, x + c.nPml) for z, x in zip(c.src_z, c.src_x)]
    recs = [(c.rec_z + c.nPml, c.rec_x + c.nPml)] * len(src)
    lam, mu, den = models.moduli(c, "true")
    obs = oracle_mod.run(g, 2, lam, mu, den, stf, src, recs, precision=32)["syn"]
    lam0, mu0, den0 = models.moduli(c, "start")
    o32 = oracle_mod.run(g, 1, lam0, mu0, den0, stf, src, recs, obs, precision=32, cond=cond)
    o64 = oracle_mod.run(g, 1, lam0, mu0, den0, stf, src, recs, obs, precision=64, cond=cond)
    j32 = oracle_mod.run(g, 0, lam0, mu0, den0, stf, src, recs, obs, precision=32,
                         cond=cond)["misfit"]
    P = fwiflow.Plan(para, ids)
    try:
        P.set_model(lam0, mu0, den0)
        P.set_stf(stf)
        P.set_obs(obs)
        P.run(1)
        J = P.misfit()
        res = P.residual()
        srcrows = P.source()
        gL, gM, gD, gS = P.gradients()
    finally:
        P.close()
    nrec, nt = len(c.rec_x), c.nSteps
    res = res.reshape(len(ids), nrec, nt)
    out = {"misfit32": abs(J - j32) / abs(j32),
           "res32": max(rel(res[s], o32["res"][s]) for s in ids),
           "res64": max(rel(res[s], o64["res"][s]) for s in ids),
           "lam32": rel(gL, o32["grad_lambda"]), "mu32": rel(gM, o32["grad_mu"]),
           "den32": rel(gD, o32["grad_den"]),
           "lam64": rel(gL, o64["grad_lambda"]), "mu64": rel(gM, o64["grad_mu"]),
           "den64": rel(gD, o64["grad_den"])}
    if cond is not None and cond["if_src_update"]:
        out["src32"] = max(rel(srcrows[s], o32["src_updated"][s]) for s in ids)
    say(variant, {k: f"{v:.2e}" for k, v in out.items()})
    for k, v in out.items():
        tol = SEIS_TOL if k.startswith(("misfit", "res", "src")) else GRAD_TOL
        assert v <= tol, (k, v, tol)


def test_scratch_files_and_unconditioned_path_agree(oracle_mod, workspace):
    """A parameter file with scratch_dir_name takes the shot-by-shot residual
    path (it keeps the conditioned data): its residual and gradients equal
    the default path's bit for bit, its misfit to summation order, and the
    scratch files (libCUFD.cu:493-511) hold the residual, the windowed data and
    the source rows."""
    out = {}
    for scratch in (False, True):
        ws = os.path.join(workspace, "scr" if scratch else "plain")
        c, para, stf, _ = _cond_setup(ws, scratch=scratch)
        ids = list(range(len(c.src_x)))
        lam, mu, den = models.moduli(c, "true")
        fwiflow.fwi_obs_op(lam, mu, den, stf, 0, ids, para)
        lam0, mu0, den0 = models.moduli(c, "start")
        P = fwiflow.Plan(para, ids)
        try:
            P.set_model(lam0, mu0, den0)
            P.set_stf(stf)
            P.load_obs()
            P.run(1)
            out[scratch] = (P.misfit(), P.residual(), P.gradients(), P.data())
        finally:
            P.close()
        if scratch:  # fwiflow_cufd writes the scratch files after a gradient run
            fwiflow.fwi_op_grad(lam0, mu0, den0, stf, 0, ids, para)
    (J0, r0, g0, d0), (J1, r1, g1, d1) = out[False], out[True]
    assert J1 == pytest.approx(J0, rel=1e-12)
    assert np.array_equal(r0, r1)
    for a, b in zip(g0, g1):
        assert np.array_equal(a, b)
    nrec, nt = len(c.rec_x), c.nSteps
    scr = os.path.join(workspace, "scr", "scr")
    w2 = fwiflow.host_tables(para, c.NZ - c.nPad, c.NX, c.nSteps)[2]
    for s in ids:
        res = np.fromfile(os.path.join(scr, f"Residual_Shot{s}.bin"), dtype=np.float32)
        assert np.array_equal(res, r1.reshape(len(ids), -1)[s])
        syn = np.fromfile(os.path.join(scr, f"Syn_Shot{s}.bin"), dtype=np.float32)
        want = (d1.reshape(len(ids), nrec, nt)[s] * w2[None, :].astype(np.float32))
        assert np.array_equal(syn.reshape(nrec, nt), want.astype(np.float32))
        assert os.path.getsize(os.path.join(scr, f"CondObs_Shot{s}.bin")) == 4 * nrec * nt
        src = np.fromfile(os.path.join(scr, f"src_updated{s}.bin"), dtype=np.float32)
        assert src.size == nt and np.abs(src).max() > 0


def test_irregular_survey_parity(oracle_mod, workspace):
    """Shots with their own receiver sets (Src_Rec.cu:94-117 reads z_rec /
    x_rec per shot): a vertical line of receivers plus surface ones, a short
    set at another depth, and two receivers in one cell.  Receivers off one
    row take the separate residual-injection launch; two in one cell add
    both residuals.  Misfit and gradients against the oracle."""
    import json
    c = models.case_small(nz=40, nx=60, nSteps=300, nshots=3)
    para, stf = models.write_case(c, workspace)
    d = json.loads(open(para).read())
    sv = json.loads(open(d["survey_fname"]).read())
    recs = [
        (list(range(2, 30)) + [1] * 20, [40] * 28 + list(range(5, 45, 2))),
        ([6] * 11, list(range(10, 54, 4))),
        ([3, 3, 3, 12], [20, 20, 33, 33]),
    ]
    for s, (zr, xr) in enumerate(recs):
        sv[f"shot{s}"].update(z_rec=zr, x_rec=xr, nrec=len(zr))
    open(d["survey_fname"], "w").write(json.dumps(sv))
    ids = [0, 1, 2]
    lam, mu, den = models.moduli(c, "true")
    say("irregular survey", [len(r[0]) for r in recs])
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, ids, para)
    lam0, mu0, den0 = models.moduli(c, "start")
    J = fwiflow.fwi_op(lam0, mu0, den0, stf, 0, ids, para)
    gL, gM, gD, gS = fwiflow.fwi_op_grad(lam0, mu0, den0, stf, 0, ids, para)
    Jo = oracle_mod.cufd(lam0, mu0, den0, stf, 0, ids, para, precision=32)[0]
    r = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=32)
    out = {"misfit32": abs(J - Jo) / abs(Jo), "lam32": rel(gL, r[1]), "mu32": rel(gM, r[2]),
           "den32": rel(gD, r[3])}
    say({k: f"{v:.2e}" for k, v in out.items()})
    for k, v in out.items():
        assert v <= (SEIS_TOL if k.startswith("misfit") else GRAD_TOL), (k, v)


def test_autotuner_measures_caches_and_applies(workspace):
    """fwiflow.tune.autotune times launch shapes around the automatic one on
    a shortened record of the caller's problem, caches the fastest per
    (grid, shot count), and a new plan given that shape reproduces the
    automatic plan's data bit for bit and gradients to the summation order
    of a launch's shots."""
    from fwiflow import tune

    c, para, stf, ids, m0 = _grad_setup(workspace, nshots=4)
    path = os.path.join(workspace, "launch.json")
    e = tune.autotune(para, ids, *m0, stf, steps=60, repeats=1, path=path, log=say)
    say("tuned", {k: v for k, v in e.items() if k != "candidates"})
    assert e["tile_x"] in (4, 2) and 1 <= e["shots_fwd"] <= 4 and 1 <= e["shots_bwd"] <= 4
    assert len(e["candidates"]) >= 2 and e["seconds"] == min(e["candidates"].values())
    assert tune.lookup(para, len(ids), path) == e
    d0, g0, J0, _ = _run_plan(para, ids, *m0, stf)
    with fwiflow.Plan(para, ids) as P:
        assert tune.apply_tuned(P, para, len(ids), path) == e
        assert P.launch() == (e["tile_x"], e["shots_fwd"], e["shots_bwd"])
        P.set_model(*m0)
        P.set_stf(stf)
        P.load_obs()
        P.run(2)
        assert np.array_equal(P.data(), d0)
        P.run(1)
        for a, b in zip(P.gradients(), g0):
            assert rel(a, b) < 1e-5


def test_reference_gradtest_taylor(workspace):
    """The reference's own gradient test (deps/CustomOps/FWI/gradtest.jl:
    400x400, nSteps 2001, one central source, a 13x13 receiver mesh, the
    misfit as a function of mu at the mu of Vp 3200, a positive random
    direction off the C-PML; restated in tests/golden/make_gradtest.py) on
    the GPU.  The reference only plots |f(m + g v) - f(m)| = O(g) and
    |f(m + g v) - f(m) - g <v, grad f>| = O(g^2) (gradtest.jl:137-171); here
    both slopes are asserted, and <v, grad f>, the differences and the
    remainders are held to the oracle's (FP32 reference arithmetic for the
    directional derivative, FP64 for the differences, which the GPU sums in
    double)."""
    sys.path.insert(0, GOLDEN)
    import make_gradtest as G

    fx = np.load(os.path.join(GOLDEN, "gradtest.npz"))
    para, lam, mu, den, mu2, stf, v = G.setup(workspace)
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, [0], para)
    with fwiflow.Plan(para, [0]) as P:
        P.set_stf(stf)
        P.load_obs()

        def f(m, calc=0):
            P.set_model(lam, m, den)
            P.run(calc)
            return P.misfit()

        y0 = f(mu2, 1)
        vg = float(np.sum(v * P.gradients()[1]))
        s = np.array([f(mu2 + g * v) - y0 for g in fx["gamma"]])
    w = s - fx["gamma"] * vg
    say("gradtest y0", y0, "oracle", float(fx["y64"]), "<v,g>", vg, "oracle FP32",
        float(fx["vg32"]), "s", s, "oracle FP64", fx["s64"], "w", w, "oracle FP64", fx["w64"])
    assert abs(y0 - fx["y64"]) <= 1e-4 * fx["y64"]
    assert abs(vg - fx["vg32"]) <= 1e-3 * abs(fx["vg32"])
    assert np.all(np.abs(s[:2] - fx["s64"][:2]) <= 1e-4 * np.abs(fx["s64"][:2]))
    slope_s = np.log10(s[0] / s[1])
    slope_w = np.log10(w[0] / w[1])
    assert abs(slope_s - 1.0) < 0.05 and abs(slope_w - 2.0) < 0.1, (slope_s, slope_w)
    assert np.all(np.abs(w[:2] - fx["w64"][:2]) <= 0.02 * np.abs(fx["w64"][:2]))


@pytest.mark.parametrize("mode", ["sweeps_of_one_shot", "checkpoints_7"])
def test_conditioning_with_sweeps_and_checkpoints(oracle_mod, workspace, mode):
    """All three conditioning stages (windows and weights, band-pass, source
    update) in a gradient run split into sweeps of one shot, or with
    checkpoints every 7 steps: residual, updated source rows and misfit equal
    the single-sweep run's (the stages work shot by shot); gradients against
    the oracle (its scheme 2 for the checkpoints: the forward segments are
    recomputed with the original source rows, the reversal subtracts the
    updated ones, libCUFD.cu:326-329,390-394)."""
    c, para, stf, cond = _cond_setup(workspace, if_win=True, filter_para=[2.0, 5.0, 25.0, 40.0],
                                     if_src_update=True)
    ids = list(range(len(c.src_x)))
    g = oracle_mod.geom(c.NZ, c.NX, c.nPml, c.nPad, c.nSteps, c.dz, c.dx, c.dt, c.f0)
    src = [(z + c.nPml, x + c.nPml) for z, x in zip(c.src_z, c.src_x)]
    recs = [(c.rec_z + c.nPml, c.rec_x + c.nPml)] * len(src)
    obs = oracle_mod.run(g, 2, *models.moduli(c, "true"), stf, src, recs, precision=32)["syn"]
    m0 = models.moduli(c, "start")

    # sweeps of one shot need backward groups of one shot (the whole ring of a
    # group must fit, else the run switches to checkpoints)
    group = {"FWIFLOW_SHOT_GROUP": "1"} if mode == "sweeps_of_one_shot" else {}

    def run(budget=None, k=None):
        with env(**group), fwiflow.Plan(para, ids) as P:
            P.set_model(*m0)
            P.set_stf(stf)
            P.set_obs(obs)
            if budget is not None:
                P.set_memory_budget(budget)
            if k is not None:
                P.set_checkpoint(k)
            P.run(1)
            return P.misfit(), P.residual(), P.source(), P.gradients(), P.memory_plan()

    J0, r0, s0, g0, p0 = run()
    if mode == "sweeps_of_one_shot":
        J1, r1, s1, g1, p1 = run(budget=1.2 * p0[2] / len(ids))
        assert p1[0] == 0 and p1[1] == 1
        o = oracle_mod.run(g, 1, *m0, stf, src, recs, obs, precision=32, cond=cond)
    else:
        J1, r1, s1, g1, p1 = run(k=7)
        assert p1[0] == 7
        o = oracle_mod.run(g, 1, *m0, stf, src, recs, obs, precision=32, cond=cond, scheme=2,
                           ckpt_every=7)
    say(mode, "plan", p1, "misfit", J1, J0)
    assert np.array_equal(r1, r0) and np.array_equal(s1, s0)
    assert J1 == pytest.approx(J0, rel=1e-12)
    for name, a, b in zip(("lam", "mu", "den"), g1[:3], (o["grad_lambda"], o["grad_mu"],
                                                          o["grad_den"])):
        e = rel(a, b)
        say(mode, name, "rel to oracle", e)
        assert e < GRAD_TOL
