"""Learned preconditioner: the network, its training, and the inversion driver
that uses it with a plain fallback.

No reference counterpart: FwiFlow's plain method is L-BFGS-B on cp
(docs/codes/src_fwi_coupled/main_mar_fwi.jl:184-187).  Here a 4-layer 3x3
convolutional network (HIP kernels behind ``fwiflow_cnn_*``: 4 -> 32 -> 32 ->
32 -> 3 channels, GELU, FP32 weights, BF16 activations) maps the raw cp / cs /
rho gradients plus the forward illumination to a search direction.  It is
trained on the (gradient, illumination, accepted update) pairs that the plain
driver (``inversion.lbfgs``) logs, and used by ``learned_descent``: each
iteration tries the network's direction with a backtracking Armijo line search
and falls back to the illumination-scaled gradient whenever the learned step
is not a descent direction or the line search rejects it; the operator's own
misfit judges every step.

Feature scaling (every map is scale-free, so one network serves any model):
gradient planes are divided by their RMS over the mask, the illumination
plane is sqrt(I / max I); targets are the accepted update divided by its RMS
over the mask.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from ._lib import check, ptr
from .inversion import CP_BOUNDS

C = _lib.C


class CNN:
    """The preconditioner network on one height x width grid (device-resident)."""

    def __init__(self, height, width, gpu_id=0, seed=0):
        L = _lib.load()
        self._L = L
        self._h = C.c_void_p()
        check(L.fwiflow_cnn_create(C.byref(self._h), int(gpu_id), int(height), int(width),
                                   int(seed)))
        self.height, self.width = int(height), int(width)
        n = C.c_int(0)
        check(L.fwiflow_cnn_num_params(self._h, C.byref(n)))
        self.num_params = int(n.value)

    def close(self):
        if self._h:
            self._L.fwiflow_cnn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _f32(self, a, planes):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.size != planes * self.height * self.width:
            raise ValueError(f"expected {planes} planes of {self.height}x{self.width}")
        return a

    def get_params(self):
        out = np.zeros(self.num_params, dtype=np.float32)
        check(self._L.fwiflow_cnn_get_params(self._h, ptr(out, C.c_float)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params:
            raise ValueError("wrong parameter count")
        check(self._L.fwiflow_cnn_set_params(self._h, ptr(p, C.c_float)))

    def forward(self, x):
        x = self._f32(x, 4)
        out = np.zeros((3, self.height, self.width), dtype=np.float32)
        check(self._L.fwiflow_cnn_forward(self._h, ptr(x, C.c_float), ptr(out, C.c_float)))
        return out

    def gradients(self, x, target, mask):
        x, t, m = self._f32(x, 4), self._f32(target, 3), self._f32(mask, 1)
        loss = np.zeros(1)
        g = np.zeros(self.num_params, dtype=np.float32)
        check(self._L.fwiflow_cnn_gradients(self._h, ptr(x, C.c_float), ptr(t, C.c_float),
                                            ptr(m, C.c_float), ptr(loss, C.c_double),
                                            ptr(g, C.c_float)))
        return float(loss[0]), g

    def train_step(self, x, target, mask, lr):
        x, t, m = self._f32(x, 4), self._f32(target, 3), self._f32(mask, 1)
        loss = np.zeros(1)
        check(self._L.fwiflow_cnn_train_step(self._h, ptr(x, C.c_float), ptr(t, C.c_float),
                                             ptr(m, C.c_float), float(lr),
                                             ptr(loss, C.c_double)))
        return float(loss[0])


def _rms(a, mask):
    v = a[mask != 0]
    r = float(np.sqrt(np.mean(v * v))) if v.size else 0.0
    return r if r > 0 else 1.0


def features(grads, illum, mask):
    """Network input [4][H][W]: (cp, cs, rho) gradients over their RMS in the
    mask, and sqrt(illumination / max)."""
    out = np.zeros((4,) + mask.shape, dtype=np.float32)
    for k, g in enumerate(grads):
        g = np.asarray(g, dtype=np.float64).reshape(mask.shape) * mask
        out[k] = g / _rms(g, mask)
    il = np.asarray(illum, dtype=np.float64).reshape(mask.shape)
    out[3] = np.sqrt(np.maximum(il, 0.0) / max(il.max(), 1e-30))
    return out


def target_from_update(update, mask):
    """Target [3][H][W]: the accepted cp update over its RMS in the mask (cs and
    rho planes zero: the plain method inverts cp only)."""
    t = np.zeros((3,) + mask.shape, dtype=np.float32)
    u = np.asarray(update, dtype=np.float64).reshape(mask.shape) * mask
    t[0] = u / _rms(u, mask)
    return t


def diagonal_scale(gradient, extras, mask, net=None, smooth=3.0, clip=(0.2, 5.0),
                   illum_eps=1e-3, illum_power=0.0):
    """A diagonal preconditioner for ``inversion.lbfgs(scale=...)`` from one
    evaluation at the starting model: with a network, the ratio of the
    magnitudes of its predicted update and of the gradient (both over their
    RMS in the mask) cell by cell, the diagonal inverse Hessian it has
    learned; without one (the control), 1 / sqrt(illumination / max + eps),
    the pseudo-Hessian scaling.  ``illum_power`` > 0 multiplies the network's
    diagonal by the illumination one to that power.  Gaussian-smoothed
    (``smooth`` cells), normalised to mean 1 in the mask and clipped to
    ``clip``; 1 outside it."""
    from scipy.ndimage import gaussian_filter

    m2 = np.asarray(mask)
    il = np.asarray(extras.get("illumination", np.ones(m2.shape)),
                    dtype=np.float64).reshape(m2.shape)
    s_il = 1.0 / np.sqrt(np.maximum(il, 0.0) / max(il.max(), 1e-30) + illum_eps)
    if net is None:
        s = s_il
    else:
        grads = extras.get("gradients")
        g = np.asarray(gradient, dtype=np.float64).reshape(m2.shape) * m2
        gq = (g,) + tuple(grads[1:] if grads is not None else (np.zeros(m2.shape),) * 2)
        d = net.forward(features(gq, il, m2)).astype(np.float64)[0] * m2
        d = np.abs(d) / _rms(d, m2)
        q = np.abs(g) / _rms(g, m2)
        s = d / (q + 0.1)
        if illum_power:
            s = s * s_il ** illum_power
    if smooth:
        s = gaussian_filter(s, smooth)
    inside = m2 != 0
    s = s / max(float(np.mean(s[inside])), 1e-30)
    s = np.clip(s, *clip)
    s[~inside] = 1.0
    return s


def train(net, samples, mask, epochs=10, lr=1e-3, seed=0, log=None):
    """Adam on (features, target) pairs in random order; returns the mean loss
    of each epoch."""
    rng = np.random.default_rng(seed)
    m = np.asarray(mask, dtype=np.float32)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = 0.0
        for i in order:
            x, t = samples[i]
            tot += net.train_step(x, t, m, lr)
        hist.append(tot / max(1, len(samples)))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4f}")
    return hist


def learned_descent(objective, x0, net, mask, bounds=CP_BOUNDS, maxiter=100,
                    target_misfit=None, step0=20.0, c_armijo=1e-4, max_backtracks=3,
                    illum_eps=1e-3):
    """Descent on cp with the network's direction, plain fallback.

    ``objective(x) -> (misfit, dJ/dcp, extras)`` with extras ``illumination``
    and ``gradients`` (cp, cs, rho), e.g. ``inversion.PlanObjective``.
    ``step0``: first trial step, RMS model change in m/s over the mask.  Every
    trial is a full operator run (its gradient is the next iteration's).
    Returns dict: x, misfit, iterations, evaluations, seconds, history
    (seconds, misfit, direction used) per accepted iteration.
    """
    shape = np.asarray(x0).shape
    x = np.asarray(x0, dtype=np.float64).ravel().copy()
    mflat = np.asarray(mask, dtype=np.float64).ravel()
    lo, hi = bounds
    t0 = time.perf_counter()
    nev = 0

    def evaluate(v):
        nonlocal nev
        nev += 1
        J, g, ex = objective(v)
        return float(J), np.asarray(g, dtype=np.float64).ravel(), ex

    J, g, ex = evaluate(x)
    hist = []
    step = float(step0)

    def try_direction(d):
        """Armijo backtracking along unit-RMS direction d, each new trial at the
        minimiser of the quadratic through J(0), its slope and the rejected
        J(alpha), kept within [0.1, 0.5] alpha; returns the accepted
        (x, J, g, ex, alpha) or None."""
        slope = float(np.dot(g, d))
        if not slope < 0:
            return None
        alpha = step
        for _ in range(max_backtracks + 1):
            xn = np.clip(x + alpha * d, lo, hi)
            Jn, gn, exn = evaluate(xn)
            if Jn <= J + c_armijo * float(np.dot(g, xn - x)) and Jn < J:
                return xn, Jn, gn, exn, alpha
            curv = Jn - J - slope * alpha
            a_q = -slope * alpha * alpha / (2.0 * curv) if curv > 0 else 0.5 * alpha
            alpha = min(max(a_q, 0.1 * alpha), 0.5 * alpha)
        return None

    for _ in range(maxiter):
        grads = ex.get("gradients", (g.reshape(shape), np.zeros(shape), np.zeros(shape)))
        il = ex.get("illumination", np.ones(shape))
        d = net.forward(features(grads, il, np.asarray(mask))).astype(np.float64)[0].ravel()
        d = d * mflat
        d /= _rms(d.reshape(shape), np.asarray(mask))
        used = "learned"
        res = try_direction(d)
        if res is None:
            ilf = np.asarray(il, dtype=np.float64).ravel()
            dp = -g / (ilf + illum_eps * max(ilf.max(), 1e-30)) * mflat
            dp /= _rms(dp.reshape(shape), np.asarray(mask))
            used = "fallback"
            res = try_direction(dp)
        if res is None and step > 1e-3 * step0:
            step *= 0.05  # both directions rejected at this scale: retry much shorter
            continue
        if res is None:
            break
        x, J, g, ex, alpha = res
        step = alpha * 2.0
        hist.append((time.perf_counter() - t0, J, used))
        if target_misfit is not None and J <= target_misfit:
            break
    return {"x": x.reshape(shape), "misfit": J, "iterations": len(hist), "evaluations": nev,
            "seconds": time.perf_counter() - t0, "history": hist}


def learned_lbfgs(objective, x0, mask, net=None, bounds=CP_BOUNDS, maxiter=100, memory=5,
                  target_misfit=None, first_step_rms=10.0, c_armijo=1e-4, max_backtracks=4,
                  illum_eps=1e-3, on_iter=None, on_pair=None, max_seconds=None):
    """L-BFGS on cp with the network as the initial inverse Hessian.

    The two-loop recursion (Nocedal & Wright, Alg. 7.4) applies H0 to the
    partly corrected gradient q: with a network, H0 q = gamma * RMS(q) *
    (-net(features(q)))'s unit-RMS direction (the network predicts an update,
    i.e. -H g); without one (``net=None``, the control), H0 q = gamma * q /
    (illumination + eps), the classic pseudo-Hessian scaling.  gamma comes
    from the latest curvature pair (s'y / y'y, scaled to H0's RMS response),
    or makes the first step ``first_step_rms`` m/s RMS.  The line search is
    backtracking Armijo from alpha = 1 with quadratic interpolation and no
    curvature condition (the reference's gradient is approximate, so a strong
    Wolfe search stalls); pairs with s'y <= 0 are skipped, and a direction
    that is not a descent direction falls back to H0 g alone, then to the
    illumination-scaled gradient.  When every candidate fails, the curvature
    memory is cleared and the illumination-scaled gradient is tried once more
    at a tenth of the step; only then does the run stop.  ``on_iter(k, J,
    used, alpha, evaluations)`` is called after each accepted step and
    ``on_pair(gradient, extras, update, misfit)`` with the training pair of
    that step, as ``inversion.lbfgs`` logs them.  ``max_seconds`` stops the
    run after the first step that ends past that wall time.  Returns the dict
    of ``learned_descent``.
    """
    shape = np.asarray(x0).shape
    m2 = np.asarray(mask)
    mflat = m2.astype(np.float64).ravel()
    x = np.clip(np.asarray(x0, dtype=np.float64).ravel(), *bounds)
    lo, hi = bounds
    t0 = time.perf_counter()
    nev = 0

    def evaluate(v):
        nonlocal nev
        nev += 1
        J, g, ex = objective(v)
        return float(J), np.asarray(g, dtype=np.float64).ravel() * mflat, ex

    def rms(v):
        return _rms(v.reshape(shape), m2)

    J, g, ex = evaluate(x)
    S, Y = [], []
    hist = []

    def h0(q, ex, gamma):
        il = np.asarray(ex.get("illumination", np.ones(shape)), dtype=np.float64).ravel()
        if net is None:
            return gamma * q / (il + illum_eps * max(il.max(), 1e-30)) * mflat
        grads = ex.get("gradients")
        gq = (q.reshape(shape),) + tuple(grads[1:] if grads is not None
                                         else (np.zeros(shape), np.zeros(shape)))
        d = net.forward(features(gq, il.reshape(shape), m2)).astype(np.float64)[0].ravel()
        d = d * mflat
        d /= rms(d)
        return -gamma * rms(q) * d  # H0 q: the network's update is -H q

    def gamma_now(ex):
        if S:
            s, y = S[-1], Y[-1]
            # scale H0 so that H0 y has the size of s along y
            r0 = h0(y, ex, 1.0)
            den = float(np.dot(y, r0))
            return float(np.dot(s, y)) / den if den > 0 else 1.0
        r0 = h0(g, ex, 1.0)
        return first_step_rms / max(rms(r0), 1e-30)

    for _ in range(maxiter):
        gamma = gamma_now(ex)
        q = g.copy()
        alphas = []
        for s, y in reversed(list(zip(S, Y))):
            rho = 1.0 / float(np.dot(y, s))
            a = rho * float(np.dot(s, q))
            alphas.append((a, rho, s, y))
            q -= a * y
        r = h0(q, ex, gamma)
        for a, rho, s, y in reversed(alphas):
            b = rho * float(np.dot(y, r))
            r += s * (a - b)
        candidates = [(-r, "lbfgs")]
        if S:
            candidates.append((-h0(g, ex, gamma), "h0"))
        il = np.asarray(ex.get("illumination", np.ones(shape)), dtype=np.float64).ravel()
        dp = -g / (il + illum_eps * max(il.max(), 1e-30)) * mflat
        dp *= first_step_rms / rms(dp)
        candidates += [(dp, "fallback"), (0.1 * dp, "restart")]
        res = None
        for d, used in candidates:
            if used == "restart":
                S.clear()
                Y.clear()
            slope = float(np.dot(g, d))
            if not slope < 0:
                continue
            alpha = 1.0
            for _ in range(max_backtracks + 1):
                xn = np.clip(x + alpha * d, lo, hi)
                Jn, gn, exn = evaluate(xn)
                if Jn < J and Jn <= J + c_armijo * float(np.dot(g, xn - x)):
                    res = (xn, Jn, gn, exn, used, alpha)
                    break
                curv = Jn - J - slope * alpha
                a_q = -slope * alpha * alpha / (2.0 * curv) if curv > 0 else 0.5 * alpha
                alpha = min(max(a_q, 0.1 * alpha), 0.5 * alpha)
            if res is not None:
                break
        if res is None:
            break
        xn, Jn, gn, exn, used, alpha = res
        if on_pair is not None:
            on_pair(g, ex, xn - x, J)
        s, y = xn - x, gn - g
        if float(np.dot(s, y)) > 1e-12 * float(np.dot(y, y)):
            S.append(s)
            Y.append(y)
            if len(S) > memory:
                S.pop(0)
                Y.pop(0)
        x, J, g, ex = xn, Jn, gn, exn
        hist.append((time.perf_counter() - t0, J, used))
        if on_iter is not None:
            on_iter(len(hist), J, used, alpha, nev)
        if target_misfit is not None and J <= target_misfit:
            break
        if max_seconds is not None and time.perf_counter() - t0 > max_seconds:
            break
    return {"x": x.reshape(shape), "misfit": J, "iterations": len(hist), "evaluations": nev,
            "seconds": time.perf_counter() - t0, "history": hist}
