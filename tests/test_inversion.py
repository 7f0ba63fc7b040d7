"""The inversion driver's host logic on CPU objectives (no GPU): bounds,
convergence, evaluation/iteration logs and the (gradient, illumination,
update) pairs."""
import json
import os

import numpy as np
import pytest

from fwiflow import inversion


def _quadratic(n=40, seed=0):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) / np.sqrt(n) + 2.0 * np.eye(n)
    xs = rng.uniform(1000.0, 6000.0, size=n)  # partly outside the (1500, 5500) bounds
    calls = []

    def f(x):
        r = A @ (x - xs)
        calls.append(x.copy())
        return 0.5 * float(r @ r), A.T @ r, {"illumination": np.abs(x)}
    return f, A, xs, calls


def test_bounded_quadratic_converges_to_projected_minimum(tmp_path):
    f, A, xs, _ = _quadratic()
    x0 = np.full(xs.size, 3000.0)
    r = inversion.lbfgs(f, x0, log_dir=str(tmp_path), maxiter=500, ftol=1e-14, gtol=1e-10)
    x = r["x"]
    assert np.all(x >= 1500.0) and np.all(x <= 5500.0)
    # KKT: zero gradient on free variables, pushing outward on active bounds
    g = A.T @ (A @ (x - xs))
    free = (x > 1500.0 + 1e-6) & (x < 5500.0 - 1e-6)
    assert np.max(np.abs(g[free])) < 1e-4 * np.max(np.abs(A.T @ A @ xs))
    assert np.all(g[x <= 1500.0 + 1e-6] >= -1e-6) and np.all(g[x >= 5500.0 - 1e-6] <= 1e-6)
    assert r["misfit"] < f(x0)[0]


def test_pairs_chain_the_iterates(tmp_path):
    f, A, xs, calls = _quadratic(seed=1)
    x0 = np.full(xs.size, 2500.0)
    r = inversion.lbfgs(f, x0, log_dir=str(tmp_path), maxiter=15)
    k = r["iterations"]
    assert k > 0
    pairs = sorted(os.listdir(tmp_path / "pairs"))
    assert len(pairs) == k
    evaluated = np.array(calls)
    x = x0.copy()
    for name in pairs:
        p = np.load(tmp_path / "pairs" / name)
        J, g, ex = f(x)
        np.testing.assert_allclose(p["gradient"], g.astype(np.float32), rtol=1e-6)
        np.testing.assert_allclose(p["illumination"], ex["illumination"].astype(np.float32))
        assert float(p["misfit"]) == pytest.approx(J, rel=1e-12)
        # the update (stored in float32) leads to an evaluated point: snap to it
        y = x + p["update"].astype(np.float64)
        k_near = np.argmin(np.abs(evaluated - y).max(axis=1))
        assert np.abs(evaluated[k_near] - y).max() < 1e-3
        x = evaluated[k_near]
    np.testing.assert_array_equal(x, r["x"])
    evals = [json.loads(l) for l in open(tmp_path / "evals.jsonl")]
    iters = [json.loads(l) for l in open(tmp_path / "iters.jsonl")]
    assert len(iters) == k and len(evals) >= k + 1
    assert all(a["seconds"] <= b["seconds"] for a, b in zip(evals, evals[1:]))
    assert iters[-1]["misfit"] <= evals[0]["misfit"]


def test_target_misfit_stops_early(tmp_path):
    f, A, xs, _ = _quadratic(seed=2)
    x0 = np.full(xs.size, 2000.0)
    full = inversion.lbfgs(f, x0, maxiter=200)
    J0 = f(x0)[0]
    target = J0 - 0.5 * (J0 - full["misfit"])
    r = inversion.lbfgs(f, x0, maxiter=200, target_misfit=target)
    assert r["message"] == "target misfit reached"
    assert r["misfit"] <= target
    assert r["iterations"] < full["iterations"]


def test_reference_defaults():
    assert inversion.CP_BOUNDS == (1500.0, 5500.0)
    assert (inversion.MAXITER, inversion.FTOL, inversion.GTOL) == (100, 1e-6, 1e-6)


def test_evaluation_cache_keeps_a_restored_point():
    """Regression: a point stored again after other points must survive the
    eviction of its earlier entry (the driver's callback re-evaluates an
    evicted iterate and then reads it back)."""
    c = inversion._EvalCache(size=3)
    pts = [np.full(4, float(k)) for k in range(6)]
    for k in range(3):
        c.put(pts[k], k)
    c.put(pts[0], "again")  # refresh point 0
    c.put(pts[3], 3)        # evicts point 1, not point 0
    assert c.get(pts[0]) == "again" and c.get(pts[1]) is None
    c.put(pts[4], 4)
    c.put(pts[5], 5)        # evicts 2 then 0
    assert c.get(pts[0]) is None and c.get(pts[5]) == 5


def test_learned_lbfgs_control_converges_on_bounded_quadratic():
    """The two-loop recursion with the illumination-scaled H0 (no network)
    reaches the projected minimum of a bounded quadratic."""
    from fwiflow import precond
    f, A, xs, _ = _quadratic(n=40, seed=3)
    x0 = np.full(xs.size, 3000.0)
    mask = np.ones(xs.size)
    r = precond.learned_lbfgs(f, x0, mask, net=None, maxiter=300, memory=8)
    x = r["x"]
    assert np.all(x >= 1500.0) and np.all(x <= 5500.0)
    g = A.T @ (A @ (x - xs))
    free = (x > 1500.0 + 1e-3) & (x < 5500.0 - 1e-3)
    assert np.max(np.abs(g[free])) < 1e-3 * np.max(np.abs(A.T @ A @ xs))
    ref = inversion.lbfgs(f, x0, maxiter=500, ftol=1e-14, gtol=1e-10)["misfit"]
    assert r["misfit"] <= ref * (1 + 1e-9)


class _IdentityNet:
    """Stands in for the network: returns the negated cp gradient plane."""

    def forward(self, x):
        return np.stack([-x[0], np.zeros_like(x[0]), np.zeros_like(x[0])])


def test_learned_lbfgs_with_gradient_net_descends():
    from fwiflow import precond
    f, A, xs, _ = _quadratic(n=36, seed=4)
    f2 = lambda v: (lambda J, g, ex: (J, g, {"illumination": np.ones(v.size),  # noqa: E731
                                             "gradients": (g.reshape(6, 6),
                                                           np.zeros((6, 6)),
                                                           np.zeros((6, 6)))}))(*f(v))
    x0 = np.full((6, 6), 3000.0)
    r = precond.learned_lbfgs(f2, x0, np.ones((6, 6)), net=_IdentityNet(), maxiter=60)
    ref = inversion.lbfgs(f, x0.ravel(), maxiter=500, ftol=1e-14, gtol=1e-10)["misfit"]
    assert r["misfit"] <= ref * (1 + 1e-9)  # the bounded minimum, as scipy's L-BFGS-B finds it
    used = [h[2] for h in r["history"]]
    assert used[:10] == ["lbfgs"] * 10  # (at the bounded minimum, later steps fall back)


def test_diagonal_scale_preconditions_lbfgs():
    """inversion.lbfgs(scale=s) runs L-BFGS-B on x / s: on a diagonal
    quadratic with curvatures over 4 decades, s = 1 / sqrt(curvature) (the
    exact diagonal preconditioner) reaches the bounded minimum in far fewer
    iterations than the unscaled run, and the bounds still hold per
    variable.  precond.diagonal_scale without a network gives the
    illumination scaling: mean 1 in the mask, clipped, 1 outside."""
    from fwiflow import precond
    n = 64
    rng = np.random.default_rng(7)
    h = 10.0 ** rng.uniform(-2, 2, size=n)
    xs = rng.uniform(1000.0, 6000.0, size=n)

    def f(x):
        return 0.5 * float(np.sum(h * (x - xs) ** 2)), h * (x - xs), {}
    x0 = np.full(n, 3000.0)
    plain = inversion.lbfgs(f, x0, maxiter=500, ftol=1e-14, gtol=1e-10)
    s = 1.0 / np.sqrt(h)
    scaled = inversion.lbfgs(f, x0, maxiter=500, ftol=1e-14, gtol=1e-10, scale=s / s.mean())
    want = np.clip(xs, 1500.0, 5500.0)
    for r in (plain, scaled):
        assert np.all(r["x"] >= 1500.0) and np.all(r["x"] <= 5500.0)
    assert np.max(np.abs(scaled["x"] - want)) < 1e-3
    assert scaled["iterations"] * 3 <= plain["iterations"], (scaled["iterations"],
                                                              plain["iterations"])
    mask = np.zeros((8, 8))
    mask[2:6, 2:6] = 1
    il = np.outer(np.linspace(1, 100, 8), np.ones(8))
    d = precond.diagonal_scale(np.ones((8, 8)), {"illumination": il}, mask, smooth=0)
    assert abs(d[mask != 0].mean() - 1.0) < 0.2 and np.all(d[mask == 0] == 1.0)
    assert np.all(np.diff(d[2:6, 3]) <= 0)  # less illumination, larger steps


class _ScaledNet:
    """Stands in for the network: predicts the cp update as -w * gradient
    plane, w a fixed positive field (a known diagonal)."""

    def __init__(self, w):
        self.w = w

    def forward(self, x):
        return np.stack([-self.w * x[0], np.zeros_like(x[0]), np.zeros_like(x[0])])


def test_diagonal_scale_recovers_a_known_network_diagonal():
    """A network whose predicted update is -w * gradient implies the diagonal
    w (up to the mean-1 normalisation) wherever the gradient is well above
    the 0.1 RMS floor; with illum_power 1 it is multiplied by the
    illumination diagonal; clip bounds hold."""
    from fwiflow import precond
    n = 16
    mask = np.ones((n, n))
    w = np.outer(np.linspace(1.0, 3.0, n), np.ones(n))
    g = np.full((n, n), 5.0)  # gradient RMS 5: |q| = 1 >> 0.1
    ex = {"illumination": np.ones((n, n)), "gradients": (g, np.zeros((n, n)), np.zeros((n, n)))}
    s = precond.diagonal_scale(g, ex, mask, _ScaledNet(w), smooth=0, clip=(0.01, 100.0))
    want = w / w.mean()
    assert np.allclose(s, want, rtol=1e-6)
    il = np.outer(np.ones(n), np.linspace(1.0, 4.0, n))
    ex["illumination"] = il
    s2 = precond.diagonal_scale(g, ex, mask, _ScaledNet(w), smooth=0, clip=(0.01, 100.0),
                                illum_power=1.0)
    s_il = 1.0 / np.sqrt(il / il.max() + 1e-3)
    want2 = (w * s_il) / (w * s_il).mean()
    assert np.allclose(s2, want2, rtol=1e-6)
    s3 = precond.diagonal_scale(g, ex, mask, _ScaledNet(w), smooth=0, clip=(0.9, 1.1),
                                illum_power=1.0)
    assert s3.min() >= 0.9 and s3.max() <= 1.1
