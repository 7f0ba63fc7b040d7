"""GPU tests of the preconditioner network (HIP kernels through the C ABI)
against the numpy reference (tests/cnn_ref.py, pinned in test_cnn_ref.py).

Tolerance: the GPU stores hidden pre-activations and activations in BF16; the
reference rounds the same values to BF16, so the remaining differences are
FP32 accumulation order plus the few values that land on the other side of a
BF16 rounding boundary: 5e-3 relative L2 for outputs, 2e-2 for parameter
gradients.
"""
import os
import sys

import numpy as np
import pytest

import cnn_ref as R
import fwiflow
from fwiflow import models
from fwiflow.precond import CNN, features, learned_descent, target_from_update, train

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _data(H, W, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(4, H, W)).astype(np.float32)
    t = rng.normal(size=(3, H, W)).astype(np.float32)
    m = (rng.uniform(size=(H, W)) > 0.2).astype(np.float32)
    return x, t, m


def test_forward_and_gradients_match_numpy():
    H, W = 37, 53  # ragged 16x16 tiles on both axes
    x, t, m = _data(H, W, 0)
    with CNN(H, W, seed=7) as net:
        assert net.num_params == R.num_params()
        p = net.get_params().astype(np.float64)
        out = net.forward(x)
        ref, _ = R.forward(p, x, bf16=True)
        e_out = rel(out, ref)
        L, g = net.gradients(x, t, m)
        Lr, gr = R.loss_and_grad(p, x, t, m, bf16=True)
        e_g = rel(g, gr)
        say("cnn forward rel", e_out, "loss", L, Lr, "grad rel", e_g)
        assert e_out < 5e-3
        assert L == pytest.approx(Lr, rel=5e-3)
        assert e_g < 2e-2
        # per layer too, so that no small layer hides behind a large one
        o = 0
        for ci, co in R.LAYERS:
            n = co * ci * 9 + co
            assert rel(g[o:o + n], gr[o:o + n]) < 3e-2, (ci, co)
            o += n


def test_adam_step_and_parameter_roundtrip():
    H, W = 20, 24
    x, t, m = _data(H, W, 1)
    with CNN(H, W, seed=3) as net:
        p0 = net.get_params()
        _, g = net.gradients(x, t, m)
        lr = 1e-3
        net.train_step(x, t, m, lr)
        p1 = net.get_params()
        # first Adam step: m_hat = g, v_hat = g^2
        expect = p0 - lr * g / (np.abs(g) + 1e-8)
        np.testing.assert_allclose(p1, expect, rtol=1e-5, atol=1e-8)
        net.set_params(p0)
        np.testing.assert_array_equal(net.get_params(), p0)


def test_training_fits_a_smoothing_map():
    """The network learns to map a noise field to its 3x3 box average (a
    function it can represent exactly): the loss falls well below its start."""
    H, W = 32, 48
    rng = np.random.default_rng(5)
    samples = []
    mask = np.ones((H, W), dtype=np.float32)
    for _ in range(8):
        x = rng.normal(size=(4, H, W)).astype(np.float32)
        xp = np.pad(x[0], 1)
        box = sum(xp[i:i + H, j:j + W] for i in range(3) for j in range(3)) / 9.0
        t = np.zeros((3, H, W), dtype=np.float32)
        t[0] = box
        samples.append((x, t))
    with CNN(H, W, seed=11) as net:
        L0 = np.mean([net.gradients(x, t, mask)[0] for x, t in samples])
        hist = train(net, samples, mask, epochs=40, lr=3e-3, seed=0)
        say("cnn training loss", L0, "->", hist[-1])
        assert hist[-1] < 0.2 * L0


def test_learned_descent_with_fallback_reduces_misfit(workspace):
    """The learned driver on a 3-shot FWI problem.  Every step is judged by the
    operator's misfit.  An untrained network still gives accepted descent
    steps; a network with all parameters zero gives no direction, so every
    step takes the fallback (illumination-scaled gradient), and the misfit
    falls either way."""
    c = models.case_small(nz=40, nx=60, nSteps=400, nshots=3)
    fwi = fwiflow.make_fwi(c.nz, c.nx, c.dz, c.dx, c.nSteps, c.dt, c.src_z, c.src_x,
                           c.rec_z, c.rec_x, nPml=c.nPml, f0=c.f0, WORKSPACE=workspace)
    fwiflow.compute_observation(fwi, c.cp, c.cs, c.rho, c.stf)
    cp0 = fwiflow.padding(c.cp0, fwi.nPml, fwi.nPad)
    obj = fwiflow.PlanObjective(fwi, c.cs, c.rho, c.stf, cp_ref=c.cp0)
    try:
        J0 = obj(cp0)[0]
        with CNN(fwi.nz_pad, fwi.nx_pad, seed=1) as net:
            r = learned_descent(obj, cp0, net, fwi.mask, maxiter=5, step0=10.0)
            net.set_params(np.zeros(net.num_params, dtype=np.float32))
            rz = learned_descent(obj, cp0, net, fwi.mask, maxiter=5, step0=10.0)
    finally:
        obj.close()
    for name, res in (("untrained", r), ("zero", rz)):
        used = [h[2] for h in res["history"]]
        say("learned descent", name, J0, "->", res["misfit"], "directions", used,
            "evals", res["evaluations"])
        assert res["iterations"] >= 1 and res["misfit"] < 0.95 * J0
        assert res["x"].min() >= 1500.0 and res["x"].max() <= 5500.0
        ms = [J0] + [h[1] for h in res["history"]]
        assert all(b < a for a, b in zip(ms, ms[1:]))  # every accepted step descends
    assert all(h[2] == "fallback" for h in rz["history"])
    x = features((np.ones(cp0.shape),) * 3, np.ones(cp0.shape), fwi.mask)
    assert x.shape == (4,) + cp0.shape
    t = target_from_update(np.ones(cp0.shape), fwi.mask)
    assert t.shape == (3,) + cp0.shape and np.all(t[1:] == 0)


@pytest.mark.timeout(900)
def test_config5_learned_lbfgs_on_a_held_out_model(workspace):
    """The learned preconditioner at its own scale: the 134x384 model family
    (224x448 with the 32-cell C-PML, nt 2000, 30 shots, elastic top rows).
    Training pairs (raw gradients + illumination -> accepted update) are
    logged by the L-BFGS control driver on two training seeds; the network
    trains on them with the HIP kernels; on a held-out seed the L-BFGS driver
    with the network as its initial inverse Hessian takes steps that each
    lower the operator's misfit.  With every network parameter zero the
    network gives no direction, so the first step (no curvature pairs yet)
    takes the fallback, the illumination-scaled gradient; later steps may
    use the curvature pairs' correction alone.  The misfit falls either
    way."""
    import bench_precond as B
    from fwiflow import precond

    samples, mask = [], None
    for seed in (100, 101):
        fwi, obj, x0 = B.setup(seed, 2000, os.path.join(workspace, f"train{seed}"))
        mask = fwi.mask

        def on_pair(g, extras, update, J):
            samples.append((precond.features(extras["gradients"], extras["illumination"], mask),
                            precond.target_from_update(update, mask)))
        try:
            r = precond.learned_lbfgs(obj, x0, mask, net=None, maxiter=4, on_pair=on_pair)
        finally:
            obj.close()
        say("training seed", seed, "misfit", r["history"][0][1] if r["history"] else None,
            "->", r["misfit"], "pairs", len(samples))
    assert mask.shape == (224, 448) and len(samples) >= 4
    fwi, obj, x0 = B.setup(1100, 2000, os.path.join(workspace, "heldout"))
    try:
        J0 = obj(x0)[0]
        with CNN(224, 448, seed=0) as net:
            hist = train(net, samples, mask.astype(np.float32), epochs=3, lr=1e-3, seed=0)
            say("training loss", hist)
            assert np.isfinite(hist[-1]) and hist[-1] < hist[0]
            rl = precond.learned_lbfgs(obj, x0, fwi.mask, net=net, maxiter=4)
            net.set_params(np.zeros(net.num_params, dtype=np.float32))
            rz = precond.learned_lbfgs(obj, x0, fwi.mask, net=net, maxiter=2)
    finally:
        obj.close()
    for name, res in (("trained", rl), ("zero", rz)):
        used = [h[2] for h in res["history"]]
        ms = [J0] + [h[1] for h in res["history"]]
        say("held-out", name, J0, "->", res["misfit"], "steps", used,
            "evaluations", res["evaluations"])
        assert res["iterations"] >= 1 and res["misfit"] < J0
        assert all(b < a for a, b in zip(ms, ms[1:]))  # every accepted step descends
        assert res["x"].min() >= 1500.0 and res["x"].max() <= 5500.0
    assert rz["history"][0][2] == "fallback"
    assert all(h[2] in ("lbfgs", "h0", "fallback", "restart") for h in rl["history"])


@pytest.mark.timeout(600)
def test_config5_learned_diagonal_preconditioner_on_a_held_out_model(workspace):
    """The learned diagonal preconditioner with the exact adjoint at its own
    scale (224x448, nt 2000, 30 shots): a network trained on pairs logged on
    two training seeds gives, with the illumination diagonal, the variable
    scaling of L-BFGS-B on a held-out seed (bench_precond.py --driver diag);
    the scale is finite, within its clip range and mean 1 in the mask, and
    every accepted iterate lowers the operator's misfit within the bounds."""
    import bench_precond as B
    from fwiflow import inversion, precond

    samples, mask = [], None
    for seed in (100, 101):
        fwi, obj, x0 = B.setup(seed, 2000, os.path.join(workspace, f"train{seed}"), "exact")
        mask = fwi.mask

        def on_pair(g, extras, update, J):
            samples.append((precond.features(extras["gradients"], extras["illumination"], mask),
                            precond.target_from_update(update, mask)))
        try:
            precond.learned_lbfgs(obj, x0, mask, net=None, maxiter=4, on_pair=on_pair)
        finally:
            obj.close()
    fwi, obj, x0 = B.setup(1100, 2000, os.path.join(workspace, "heldout"), "exact")
    try:
        J0, g0, ex0 = obj(x0)
        with CNN(224, 448, seed=0) as net:
            train(net, samples, mask.astype(np.float32), epochs=3, lr=1e-3, seed=0)
            s = precond.diagonal_scale(g0, ex0, fwi.mask, net, clip=(0.1, 10.0),
                                       illum_power=1.0)
        inside = fwi.mask != 0
        assert np.all(np.isfinite(s)) and s.min() >= 0.1 and s.max() <= 10.0
        assert abs(float(s[inside].mean()) - 1.0) < 0.3 and np.all(s[~inside] == 1.0)
        r = inversion.lbfgs(obj, x0, maxiter=6, scale=s)
        rp = inversion.lbfgs(obj, x0, maxiter=6)
    finally:
        obj.close()
    ms = [J0] + [h[1] for h in r["history"]]
    say("held-out diagonal", J0, "->", r["misfit"], "evaluations", r["evaluations"],
        "plain after as many iterations", rp["misfit"])
    assert r["iterations"] >= 1 and r["misfit"] < J0
    assert all(b < a for a, b in zip(ms, ms[1:]))
    assert r["x"].min() >= 1500.0 and r["x"].max() <= 5500.0
