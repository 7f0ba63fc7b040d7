"""GPU parity on the large grids, against oracle results stored as sketches.

The oracle needs minutes per case on these grids, so tests/golden/
make_sketches.py ran it once (FP32 reference arithmetic and FP64) and stored
64-probe random-projection sketches of the seismograms and gradients
(tests/sketch.py), the misfits and the source-function gradients.  Here the
GPU computes the same case from the same seeded models and its results are
sketched with the same probes.  A sketch estimates a relative L2 distance
to within about 20%, far inside the margins below.

Tolerances as in test_gpu.py: seismograms and misfits 1e-4 relative to the
FP32 oracle, gradients 1e-3; against FP64, the contract tolerance or the
FP32 oracle's own distance from FP64 plus twice the GPU's distance from the
FP32 oracle.
"""
import os
import sys

import numpy as np
import pytest

import fwiflow
from fwiflow import models
from sketch import rel_sketch, sketch, smooth_direction

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-4
GRAD_TOL = 1e-3
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C3_NT, C3_IDS = 1500, [0, 40]
C4_NT = 3000
C4_GRIDS = [(200, 600), (300, 900), (450, 1350)]
C2_IDS = [0, 14, 29]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fixture(name):
    path = os.path.join(GOLDEN, f"sketch_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{path} not generated (tests/golden/make_sketches.py)")
    return np.load(path)


def c4_ids(nz, nx, nshots=16):
    """The shot nearest above the monitor ellipse (same draws as models.case4)
    and the first shot."""
    rng = np.random.default_rng(4)
    rng.uniform()
    cx = rng.uniform(0.35, 0.65) * nx
    src_x = np.round(np.linspace(5, nx - 6, nshots)).astype(int)
    return [int(np.argmin(np.abs(src_x - cx))), 0]


def gpu_case(case, ids, ws, gradient=True):
    """Observed data from the true model (written as the Shot files), then
    the misfit and gradients at the starting model, all on the GPU."""
    para, stf = models.write_case(case, ws)
    lam, mu, den = models.moduli(case, "true")
    say("case", case.name, "grid", case.NZ, case.NX, "nt", case.nSteps, "shots", ids)
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, ids, para)
    data = np.concatenate([np.fromfile(os.path.join(ws, "Data", f"Shot{i}.bin"), dtype=np.float32)
                           for i in ids])
    out = {"data": data}
    m0 = models.moduli(case, "start")
    with fwiflow.Plan(para, ids) as P:
        P.set_model(*m0)
        P.set_stf(stf)
        P.load_obs()
        P.run(1 if gradient else 0)
        out["misfit"] = P.misfit()
        if gradient:
            out["lam"], out["mu"], out["den"], g = P.gradients()
            out["stf"] = g[:len(ids)]
    return out, para, stf


def compare(gpu, fx, gradient=True):
    """Relative distances of the GPU results from the oracle's FP32 and FP64
    results, each asserted."""
    res = {}
    s = sketch(gpu["data"])
    res["seis"] = (rel_sketch(s, fx["seis32"]), rel_sketch(s, fx["seis64"]),
                   rel_sketch(fx["seis32"], fx["seis64"]))
    J, J32, J64 = gpu["misfit"], float(fx["misfit32"]), float(fx["misfit64"])
    res["misfit"] = (abs(J - J32) / J32, abs(J - J64) / J64, abs(J32 - J64) / J64)
    if gradient:
        for name in ("lam", "mu", "den"):
            sg = sketch(gpu[name])
            res[name] = (rel_sketch(sg, fx[f"{name}32"]), rel_sketch(sg, fx[f"{name}64"]),
                         rel_sketch(fx[f"{name}32"], fx[f"{name}64"]))
        a, b32, b64 = gpu["stf"], fx["stf32"], fx["stf64"]
        r = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))
        res["stf"] = (r(a, b32), r(a, b64), r(b32, b64))
    say({k: tuple(f"{x:.2e}" for x in v) for k, v in res.items()})
    for k, (e32, e64, spread) in res.items():
        tol = SEIS_TOL if k in ("seis", "misfit") else GRAD_TOL
        assert e32 <= tol, (k, "FP32", e32, tol)
        tol64 = max(tol, spread + 2.0 * e32 + 1e-9)
        assert e64 <= tol64, (k, "FP64", e64, tol64)
    return res


@pytest.mark.timeout(600)
def test_config3_gradient_parity(workspace):
    """500x1700 (+40-cell C-PML, 10 m, dt 1 ms, 12 Hz), shots 0 and 40 of 64,
    1500 steps: seismograms, misfit and the lambda/mu/rho/source gradients."""
    fx = fixture("c3_grad")
    c = models.case3(nSteps=C3_NT, nshots=64)
    gpu, _, _ = gpu_case(c, C3_IDS, workspace)
    compare(gpu, fx)


@pytest.mark.timeout(900)
def test_config3_gradient_against_central_differences(workspace):
    """The gradient's directional derivative sum(g * m * r) in 5 smooth
    random directions r (lambda, mu and rho perturbed together by
    m (1 +- h r), h = 1e-3) against central differences of the GPU's own
    misfit.  The reference's adjoint is approximate (material parameters
    at the updated cell, its own C-PML adjoint), so gradient and FD differ
    by a few percent; the GPU must show the oracle's gap, measured in FP64,
    and its FD quotients must match the oracle's."""
    fx = fixture("c3_fd")
    h = float(fx["h"])
    c = models.case3(nSteps=C3_NT, nshots=64)
    para, stf = models.write_case(c, workspace)
    lam, mu, den = models.moduli(c, "true")
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, C3_IDS, para)
    m0 = models.moduli(c, "start")
    margin = (c.nPml + 12, c.nPml + c.nPad)  # below the source/receiver rows, above the pad
    dirs = [[smooth_direction((c.NZ, c.NX), 1000 + 10 * k + q, 8.0, (margin, c.nPml))
             for q in range(3)] for k in range(len(fx["fd64"]))]
    with fwiflow.Plan(para, C3_IDS) as P:
        P.set_stf(stf)
        P.load_obs()
        P.set_model(*m0)
        P.run(1)
        g = P.gradients()[:3]
        gdm = np.array([sum(np.sum(gq * mq * dq) for gq, mq, dq in zip(g, m0, d)) for d in dirs])
        fd = []
        for d in dirs:
            J = []
            for sgn in (1.0, -1.0):
                P.set_model(*[mq * (1.0 + sgn * h * dq) for mq, dq in zip(m0, d)])
                P.run(0)
                J.append(P.misfit())
            fd.append((J[0] - J[1]) / (2 * h))
    fd = np.array(fd)
    fd32, fd64, gdm32, gdm64 = fx["fd32"], fx["fd64"], fx["gdm32"], fx["gdm64"]
    gap = np.abs(gdm - fd) / np.abs(fd)
    gap64 = np.abs(gdm64 - fd64) / np.abs(fd64)
    say("gdm", gdm, "oracle FP32", gdm32)
    say("fd", fd, "oracle FP32", fd32, "FP64", fd64)
    say("gradient-vs-FD gap", gap, "oracle FP64", gap64)
    assert np.all(np.abs(gdm - gdm32) <= GRAD_TOL * np.abs(gdm32))
    # the FD quotients: as close to FP64 as the FP32 oracle's, or 1e-3
    assert np.all(np.abs(fd - fd64) <= 2.0 * np.abs(fd32 - fd64) + 1e-3 * np.abs(fd64))
    # the gradient-vs-FD gap of the reference's scheme, reproduced
    assert np.all(np.abs(gap - gap64) <= 0.1 * gap64 + 2e-3)


@pytest.mark.timeout(600)
def test_config3_exact_adjoint_equals_central_differences(workspace):
    """The exact adjoint (Plan.set_adjoint("exact"), the transpose of the
    forward step) on the same grid and shots: its directional derivatives
    in 5 smooth random directions equal central differences of the GPU's
    misfit (h = 1e-3 relative, FP32, whose quotients the test above holds
    to the FP64 oracle's) to 2e-3, where the reference's gradient in the
    same directions is off by percents.  The directions vanish outside the
    interior shrunk by 2 cells: like the reference's, the gradient covers
    the interior cells, and the parameters of the first interior row and
    column also enter C-PML stencils whose forward fields the reversal does
    not re-materialise."""
    h = 1e-3
    c = models.case3(nSteps=C3_NT, nshots=64)
    para, stf = models.write_case(c, workspace)
    fwiflow.fwi_obs_op(*models.moduli(c, "true"), stf, 0, C3_IDS, para)
    m0 = models.moduli(c, "start")
    margin = (c.nPml + 12, c.nPml + c.nPad)
    hard = np.zeros((c.NZ, c.NX))
    hard[c.nPml + 2:c.NZ - c.nPad - c.nPml - 2, c.nPml + 2:c.NX - c.nPml - 2] = 1.0
    dirs = [[smooth_direction((c.NZ, c.NX), 2000 + 10 * k + q, 8.0, (margin, c.nPml)) * hard
             for q in range(3)] for k in range(5)]
    with fwiflow.Plan(para, C3_IDS) as P:
        P.set_stf(stf)
        P.load_obs()
        P.set_model(*m0)
        P.run(1)
        gref = P.gradients()[:3]
        P.set_adjoint("exact")
        P.run(1)
        g = P.gradients()[:3]
        fd = []
        for d in dirs:
            J = []
            for sgn in (1.0, -1.0):
                P.set_model(*[mq * (1.0 + sgn * h * dq) for mq, dq in zip(m0, d)])
                P.run(0)
                J.append(P.misfit())
            fd.append((J[0] - J[1]) / (2 * h))
    fd = np.array(fd)

    def dd(gr):
        return np.array([sum(np.sum(gq * mq * dq) for gq, mq, dq in zip(gr, m0, d)) for d in dirs])

    gap = np.abs(dd(g) - fd) / np.abs(fd)
    gap_ref = np.abs(dd(gref) - fd) / np.abs(fd)
    say("exact gdm", dd(g), "FD", fd, "gap", gap, "reference-adjoint gap", gap_ref)
    assert np.all(gap <= 2e-3), gap
    assert np.all(gap_ref > 10 * gap), (gap_ref, gap)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("grid", C4_GRIDS, ids=[f"{z}x{x}" for z, x in C4_GRIDS])
def test_config4_time_lapse_parity(workspace, grid):
    """Time-lapse surveys on the 12 m grids (32-cell C-PML, dt 1.5 ms,
    10 Hz, 3000 steps): baseline and monitor (-6% Vp / +2% density ellipse)
    each against the oracle, for the shot above the ellipse and the first
    shot.  The monitor's data equal the baseline's bit for bit for the first
    300 steps (before waves that met the ellipse come back) and differ from
    it afterwards by the oracle's relative amount."""
    nz, nx = grid
    ids = c4_ids(nz, nx)
    data = {}
    for monitor in (False, True):
        c = models.case4(nz, nx, monitor=monitor, nSteps=C4_NT, nshots=16)
        fx = fixture(c.name)
        assert list(fx["ids"]) == ids
        gpu, _, _ = gpu_case(c, ids, os.path.join(workspace, c.name))
        compare(gpu, fx)
        data[monitor] = gpu["data"].reshape(len(ids), len(c.rec_x), C4_NT)
    assert np.array_equal(data[False][:, :, :300], data[True][:, :, :300])
    d = float(np.linalg.norm(data[True] - data[False]) / np.linalg.norm(data[False]))
    want = float(fixture(f"c4_{nz}x{nx}_timelapse")["monitor_rel"])
    say("monitor vs baseline", d, "oracle", want)
    assert abs(d - want) <= 1e-2 * want


@pytest.mark.timeout(600)
@pytest.mark.parametrize("fuse", ["-1", "1"], ids=["automatic", "fused"])
def test_config4_surveys_in_one_job(workspace, fuse):
    """The six config-4 surveys (three grids, baseline and monitor; two shots
    each, 3000 steps) run as one job: misfits and data equal each survey's
    plan run alone bit for bit, gradients to 1e-6 (a fused launch covers
    fewer shots of each survey than the survey's own launches, and a launch
    sums its shots' gradient contributions before adding them to the
    accumulators).  Fused (every time step one launch across the grids, a
    backward group of one shot per survey by the cache rule on the summed
    bytes) and automatic (here: the surveys one after another, since each
    survey alone fills launches of both its shots)."""
    old = os.environ.get("FWIFLOW_JOB_FUSE")
    os.environ["FWIFLOW_JOB_FUSE"] = fuse
    plans, alone = [], []
    try:
        for nz, nx in C4_GRIDS:
            ids = c4_ids(nz, nx)
            for monitor in (False, True):
                c = models.case4(nz, nx, monitor=monitor, nSteps=C4_NT, nshots=16)
                ws = os.path.join(workspace, c.name)
                para, stf = models.write_case(c, ws)
                fwiflow.fwi_obs_op(*models.moduli(c, "true"), stf, 0, ids, para)
                m0 = models.moduli(c, "start")
                P = fwiflow.Plan(para, ids)
                plans.append(P)
                P.set_model(*m0)
                P.set_stf(stf)
                P.load_obs()
                P.run(1)
                alone.append((P.misfit(), P.data(), P.gradients()))
        with fwiflow.Job(plans) as J:
            J.run(1)
            assert J.fused() == (fuse == "1")
            for P, (J0, d0, g0) in zip(plans, alone):
                say("job", P.misfit(), "alone", J0)
                assert P.misfit() == J0
                assert np.array_equal(P.data(), d0)
                for a, b in zip(P.gradients(), g0):
                    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)
    finally:
        if old is None:
            os.environ.pop("FWIFLOW_JOB_FUSE", None)
        else:
            os.environ["FWIFLOW_JOB_FUSE"] = old
        for P in plans:
            P.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("water", [True, False], ids=["water_layer", "elastic_top"])
def test_config2_parity(workspace, water):
    """134x384 (+32-cell C-PML, 24 m, dt 2 ms, 8 Hz), 2000 steps, shots 0, 14
    and 29 of 30.  With the 10-row water layer (Vs = 0): seismograms and
    misfit (the reference's adjoint grows without bound over a long record
    with such a layer, so its gradient is not a parity target); with the
    top rows elastic: gradients too."""
    c = models.case2(water=water)
    fx = fixture(f"c2_{'water' if water else 'elastic'}")
    gpu, _, _ = gpu_case(c, C2_IDS, workspace, gradient=not water)
    compare(gpu, fx, gradient=not water)


@pytest.mark.timeout(600)
def test_config2_water_layer_exact_adjoint_equals_central_differences(workspace):
    """Config 2 with its 10-row water layer (Vs = 0), 2000 steps, shots 0, 14
    and 29: the reference's adjoint grows without bound there, the exact
    adjoint (the transpose of the forward step) stays the derivative of the
    misfit.  Its directional derivatives in 3 smooth random directions
    (lambda, mu, rho perturbed together by m (1 +- h r), h = 1e-3; the water's
    mu stays 0) equal central differences of the GPU's misfit to 2e-3, and
    the reference-mode gradient's do not (printed with the exact ones)."""
    h = 1e-3
    c = models.case2(water=True)
    para, stf = models.write_case(c, workspace)
    fwiflow.fwi_obs_op(*models.moduli(c, "true"), stf, 0, C2_IDS, para)
    m0 = models.moduli(c, "start")
    hard = np.zeros((c.NZ, c.NX))
    hard[c.nPml + 2:c.NZ - c.nPad - c.nPml - 2, c.nPml + 2:c.NX - c.nPml - 2] = 1.0
    margin = (c.nPml + 2, c.nPml + c.nPad)
    dirs = [[smooth_direction((c.NZ, c.NX), 3000 + 10 * k + q, 6.0, (margin, c.nPml)) * hard
             for q in range(3)] for k in range(3)]
    with fwiflow.Plan(para, C2_IDS) as P:
        P.set_stf(stf)
        P.load_obs()
        P.set_model(*m0)
        P.run(1)
        gref = P.gradients()[:3]
        P.set_adjoint("exact")
        P.run(1)
        g = P.gradients()[:3]
        fd = []
        for d in dirs:
            J = []
            for sgn in (1.0, -1.0):
                P.set_model(*[mq * (1.0 + sgn * h * dq) for mq, dq in zip(m0, d)])
                P.run(0)
                J.append(P.misfit())
            fd.append((J[0] - J[1]) / (2 * h))
    fd = np.array(fd)

    def dd(gr):
        return np.array([sum(np.sum(gq * mq * dq) for gq, mq, dq in zip(gr, m0, d)) for d in dirs])

    gap = np.abs(dd(g) - fd) / np.abs(fd)
    gap_ref = np.abs(dd(gref) - fd) / np.abs(fd)
    say("water layer: exact gdm", dd(g), "FD", fd, "gap", gap, "reference-adjoint gdm",
        dd(gref), "gap", gap_ref)
    assert all(np.all(np.isfinite(a)) for a in g)
    assert np.all(gap <= 2e-3), gap
    assert np.all(~(gap_ref <= 10 * gap)), (gap_ref, gap)  # NaN counts as off
