"""Pins of the CPU oracle (oracle/cufd_oracle.cpp) -- CPU only.

The reference cannot run here (CUDA + TensorFlow + Julia) and its only
golden output file (test/test_data.mat) is absent upstream, so the oracle is
pinned by:
  * closed forms: C-PML coefficients (Komatitsch & Martin 2007 recursive
    convolution), time window, harmonic averages, Ricker wavelet;
  * the reference's own known answers: misfit at the true model is 0
    (docs/src/tutorials/fwi.md:83-87) and the Taylor-test behaviour of
    deps/CustomOps/FWI/gradtest.jl:137-151;
  * physics: linearity in the source, time-shift equivariance, P-wave speed,
    C-PML absorption;
  * exact discrete adjoint where the reference's hand-derived adjoint is exact
    (homogeneous medium, waves inside the interior): central FD to 1e-7;
  * an independent vectorised numpy transcription (tests/numpy_ref.py) on
    tiny heterogeneous grids with the C-PML active: agreement to 1e-12;
  * the two backward schemes (boundary-ring reversal vs checkpoint
    recompute) agreeing in FP64.
"""
import os

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from fwiflow import models
from fwiflow.models import Case
from fwiflow.utils import sourceGene

import numpy_ref as NR


def _setup(oracle_mod, case, ws, precision=64, obs_model="true", obs_scale=None):
    para, stf = models.write_case(case, ws)
    lam, mu, den = models.moduli(case, obs_model)
    ids = list(range(len(case.src_x)))
    oracle_mod.cufd(lam, mu, den, stf, 2, ids, para, precision=precision)
    return para, stf, ids


# ------------------------------------------------------------------ closed forms
def test_cpml_matches_published_cpml(oracle_mod):
    """d0 = -(N+1) c ln R / (2L) (Collino & Tsogka), d(x) = d0 (x/L profile),
    K = 1 + (Kmax-1)(x/L)^N, alpha = alpha_max (1 - x/L), b = exp(-(d/K+alpha) dt),
    a = d (b-1) / (K (d + K alpha))  (Komatitsch & Martin 2007, eqs. for the
    recursive-convolution C-PML)."""
    N, nPml, dh, f0, dt = 100, 20, 10.0, 12.0, 1e-3
    out = oracle_mod.cpml(N, nPml, dh, f0, dt, precision=64)
    K, a, b, Kh, ah, bh = out
    L = nPml * dh
    d0 = -(8 + 1) * 3000.0 * np.log(8e-4) / (2 * L)
    i = np.arange(N)
    for shift, (KK, aa, bb) in ((0.0, (K, a, b)), (0.5, (Kh, ah, bh))):
        depth = np.maximum(np.maximum((nPml - i - shift) * dh, (nPml - N + i + shift) * dh), -1)
        inside = depth >= 0
        xn = np.where(inside, depth / L, 0.0)
        d = np.where(inside, d0 * (0.25 * xn + 0.75 * xn ** 8), 0.0)
        Kx = np.where(inside, 1 + xn ** 8, 1.0)
        al = np.where(inside, np.pi * f0 * (1 - xn), 0.0)
        bx = np.exp(-(d / Kx + al) * dt)
        with np.errstate(invalid="ignore", divide="ignore"):
            ax = np.where(d > 1e-6, d * (bx - 1) / (Kx * (d + Kx * al)), 0.0)
        np.testing.assert_allclose(KK, Kx, rtol=1e-12)
        np.testing.assert_allclose(bb, bx, rtol=1e-12)
        np.testing.assert_allclose(aa, ax, rtol=1e-10, atol=1e-14)
    # interior: no damping at all
    mid = slice(nPml + 1, N - nPml - 1)
    assert np.all(K[mid] == 1) and np.all(a[mid] == 0) and np.all(b[mid] == 1)
    # damping grows into the layer: a < 0 and decreasing b towards the edge
    assert np.all(a[:nPml] < 0) and np.all(np.diff(b[:nPml]) > 0)


def test_cpml_fp32_close_to_fp64(oracle_mod):
    o64 = oracle_mod.cpml(264, 32, 20.0, 10.0, 2e-3, precision=64)
    o32 = oracle_mod.cpml(264, 32, 20.0, 10.0, 2e-3, precision=32)
    np.testing.assert_allclose(o32, o64, rtol=2e-6, atol=1e-9)


def test_window_closed_form(oracle_mod):
    """utilities.cu:707-747: w = sin(pi/2 t/t1) up to t1 = ratio*T, 1, then
    cos(pi/2 (t-t2)/(T-t2)); data multiplied by w^2."""
    nt, dt, ratio = 1000, 2e-3, 0.05
    out = oracle_mod.window(np.ones((2, nt)), dt, ratio, precision=64)
    t = np.arange(nt) * dt
    T = nt * dt
    t1, t2 = ratio * T, T - ratio * T
    w = np.where(t < t1, np.sin(np.pi / 2 * t / t1),
                 np.where(t < t2, 1.0, np.cos(np.pi / 2 * (t - t2) / (T - t2))))
    np.testing.assert_allclose(out[0], w * w, atol=1e-12)
    np.testing.assert_allclose(out[1], w * w, atol=1e-12)
    # rejected window (2*offset >= T) leaves data unchanged ("Window error 2")
    same = oracle_mod.window(np.full((1, 10), 3.0), dt, 0.6, precision=64)
    assert np.all(same == 3.0)


def test_model_averages_closed_form(oracle_mod):
    rng = np.random.default_rng(0)
    NZ, NX = 20, 24
    lam = rng.uniform(5e3, 9e3, (NZ, NX))
    mu = rng.uniform(2e3, 4e3, (NZ, NX))
    mu[7, 9] = 0.0
    den = rng.uniform(1800, 2600, (NZ, NX))
    g = oracle_mod.geom(NZ, NX, 4, 0, 10, 10.0, 10.0, 1e-3, 10.0)
    amu, ba, bb, cp = oracle_mod.model_arrays(g, lam, mu, den, precision=64)
    M = mu * 1e6
    for z in range(NZ):
        for x in range(NX):
            if 2 <= z <= NZ - 3 and 2 <= x <= NX - 3:
                q = [M[z, x], M[z + 1, x], M[z, x + 1], M[z + 1, x + 1]]
                want = 0.0 if min(q) == 0 else 4.0 / sum(1.0 / v for v in q)
                assert amu[z, x] == pytest.approx(want, rel=1e-14)
                assert ba[z, x] == pytest.approx(2 / (den[z, x] + den[z + 1, x]), rel=1e-14)
                assert bb[z, x] == pytest.approx(2 / (den[z, x] + den[z, x + 1]), rel=1e-14)
            else:
                assert amu[z, x] == 0 and ba[z, x] == 1e-3 and bb[z, x] == 1e-3
    np.testing.assert_allclose(cp, np.sqrt((lam + 2 * mu) * 1e6 / den), rtol=1e-14)


def test_sourcegene_is_integrated_ricker():
    """Utils.jl:116-132: cumulative sum of the Ricker wavelet times dt."""
    f, n, dt = 10.0, 500, 2e-3
    s = sourceGene(f, n, dt)[0]
    t = np.arange(n) * dt - 1.2 / f
    r = (1 - 2 * np.pi ** 2 * f ** 2 * t ** 2) * np.exp(-np.pi ** 2 * f ** 2 * t ** 2)
    np.testing.assert_allclose(np.diff(s) / dt, r[1:], rtol=1e-10, atol=1e-12)
    assert s[0] == pytest.approx(r[0] * dt)
    # a Ricker wavelet integrates to ~0
    assert abs(s[-1]) < 1e-6


# ------------------------------------------------------------------ forward physics
def test_misfit_is_zero_at_true_model(oracle_mod, workspace):
    """docs/src/tutorials/fwi.md:83-87: misfit(true model) == 0."""
    c = models.case_small(nSteps=200)
    for prec in (32, 64):
        para, stf, ids = _setup(oracle_mod, c, os.path.join(workspace, str(prec)), precision=prec)
        lam, mu, den = models.moduli(c, "true")
        J = oracle_mod.cufd(lam, mu, den, stf, 0, ids, para, precision=prec)[0]
        if prec == 32:
            assert J == 0.0  # same arithmetic, same bits
        else:
            # the Shot files are f32, so FP64 data differ from them by rounding only
            obs = np.concatenate([np.fromfile(os.path.join(workspace, "64", "Data",
                                                           f"Shot{i}.bin"), dtype=np.float32)
                                  for i in ids])
            assert J < 1e-12 * 0.5 * np.sum(obs.astype(np.float64) ** 2)


def test_data_linear_in_source(oracle_mod):
    c = models.case_small(nSteps=250)
    g = oracle_mod.geom(c.NZ, c.NX, c.nPml, c.nPad, c.nSteps, c.dz, c.dx, c.dt, c.f0)
    lam, mu, den = models.moduli(c, "true")
    src = [(c.src_z[0] + c.nPml, c.src_x[0] + c.nPml)]
    recs = [(c.rec_z + c.nPml, c.rec_x + c.nPml)]
    d1 = oracle_mod.run(g, 2, lam, mu, den, c.stf, src, recs, precision=64)["syn"][0]
    d3 = oracle_mod.run(g, 2, lam, mu, den, -2.5 * c.stf, src, recs, precision=64)["syn"][0]
    np.testing.assert_allclose(d3, -2.5 * d1, rtol=1e-6, atol=1e-6 * np.abs(d1).max())


def test_time_shift_equivariance(oracle_mod):
    c = models.case_small(nSteps=300)
    g = oracle_mod.geom(c.NZ, c.NX, c.nPml, c.nPad, c.nSteps, c.dz, c.dx, c.dt, c.f0)
    lam, mu, den = models.moduli(c, "true")
    src = [(c.src_z[0] + c.nPml, c.src_x[0] + c.nPml)]
    recs = [(c.rec_z + c.nPml, c.rec_x + c.nPml)]
    k = 25
    s = c.stf[0]
    s_shift = np.concatenate([np.zeros(k), s[:-k]])
    d = oracle_mod.run(g, 2, lam, mu, den, s, src, recs, precision=64)["syn"][0]
    ds = oracle_mod.run(g, 2, lam, mu, den, s_shift, src, recs, precision=64)["syn"][0]
    # away from the 0.1% source taper at both ends
    np.testing.assert_allclose(ds[:, k + 5:-5], d[:, 5:-k - 5], atol=1e-5 * np.abs(d).max())


def test_p_wave_speed(oracle_mod):
    """Homogeneous acoustic medium (cs = 0): the pulse moves out at Vp; a wrong
    stiffness (lambda + 2 mu), buoyancy or dt/dx scaling shifts it."""
    nz, nx, nt, dh, dt, f0, vp = 60, 200, 500, 10.0, 1e-3, 25.0, 3000.0
    cp = np.full((nz, nx), vp)
    c = Case("homog", nz, nx, dh, dh, nt, dt, f0, 20, cp, 0 * cp, np.full((nz, nx), 2000.0),
             cp, cp, cp, np.array([30]), np.array([20]), np.full(2, 30), np.array([60, 140]),
             sourceGene(f0, nt, dt))
    g = oracle_mod.geom(c.NZ, c.NX, c.nPml, c.nPad, nt, dh, dh, dt, f0)
    lam, mu, den = models.moduli(c, "true")
    r = oracle_mod.run(g, 2, lam, mu, den, c.stf, [(30 + 20, 20 + 20)],
                       [(np.array([50, 50]), np.array([80, 160]))], precision=64)["syn"][0]
    # the stf is an integrated Ricker, so the recorded stress peaks with the pulse
    t1, t2 = np.argmax(np.abs(r[0])), np.argmax(np.abs(r[1]))
    v = (160 - 80) * dh / ((t2 - t1) * dt)
    assert v == pytest.approx(vp, rel=0.02)


def test_cpml_absorbs(oracle_mod):
    """Energy that reaches the C-PML does not come back: traces at a receiver
    match a run on a 3x larger grid (no boundary within reach) to a few %."""
    nt, dh, dt, f0 = 500, 10.0, 1e-3, 20.0

    def trace(nz, nx, sz, sx):
        cp = np.full((nz, nx), 2500.0)
        c = Case("h", nz, nx, dh, dh, nt, dt, f0, 20, cp, cp / np.sqrt(3), np.full((nz, nx), 2000.0),
                 cp, cp, cp, np.array([sz]), np.array([sx]), np.array([sz]), np.array([sx + 10]),
                 sourceGene(f0, nt, dt))
        g = oracle_mod.geom(c.NZ, c.NX, c.nPml, c.nPad, nt, dh, dh, dt, f0)
        lam, mu, den = models.moduli(c, "true")
        return oracle_mod.run(g, 2, lam, mu, den, c.stf, [(sz + 20, sx + 20)],
                              [(np.array([sz + 20]), np.array([sx + 30]))], precision=64)["syn"][0][0]

    small = trace(60, 60, 30, 25)
    big = trace(300, 300, 150, 145)
    err = np.linalg.norm(small - big) / np.linalg.norm(big)
    assert err < 0.05


def _mirror_traces(oracle_mod, nPml=20, n=120, nt=300, zs=40):
    """Homogeneous elastic model (nPad 0), the source on the x-mirror axis
    x = NX/2, receivers at the mirrored columns nPml + 3 and NX - nPml - 3;
    FP64.  Returns the two traces."""
    NZ = NX = n + 2 * nPml
    dh, dt, f0 = 10.0, 1e-3, 25.0
    g = oracle_mod.geom(NZ, NX, nPml, 0, nt, dh, dh, dt, f0)
    vp = np.full((NZ, NX), 3000.0)
    vs, rho = vp / np.sqrt(3.0), np.full((NZ, NX), 2000.0)
    lam, mu = rho * (vp ** 2 - 2 * vs ** 2) / 1e6, rho * vs ** 2 / 1e6
    xr = np.array([nPml + 3, NX - nPml - 3])
    r = oracle_mod.run(g, 2, lam, mu, rho, sourceGene(f0, nt, dt), [(zs, NX // 2)],
                       [(np.full(2, zs), xr)], precision=64)
    return r["syn"][0]


def test_velocity_x_gate_is_the_one_early_mirror_asymmetry(oracle_mod):
    """The x-PML profiles (utilities.cu:242-358) are mirror images about
    x = NX/2 (integer points i <-> NX-i, half points i+1/2 <-> NX-1-i+1/2),
    and so are el_stress's x-gates (x < nPml || x > nx-nPml-1).  el_velocity's
    right gate, x > nx-nPml (el_velocity.cu:56,71), is one column narrower:
    it leaves out the first half-point column of the right C-PML (depth
    dh/2, a_x_half != 0).  With the source on the mirror axis, the traces at
    mirrored receivers agree bit for bit until waves from the grid's outer
    edge arrive (the update range [2, nx-3] is not mirror symmetric there)
    when el_velocity takes el_stress's gate, and differ by 1e-3 of the trace
    maximum once the wave has met the right C-PML with the reference's gate.
    This separates the two gates and pins the profiles' left-right placement
    and the staggering of the x-PML terms."""
    d = _mirror_traces(oracle_mod)
    with oracle_mod.mutation("velocity_gate_of_stress"):
        m = _mirror_traces(oracle_mod)
    # outer-edge influence reaches the receivers after step ~320
    assert np.array_equal(m[0], m[1])
    assert np.abs(d[0] - d[1]).max() > 1e-3 * np.abs(d[0]).max()


def _pml_fd_gap(oracle_mod, s, mutation=None):
    """Homogeneous elastic model (Vp 3000, Vs = Vp/sqrt(3), rho 2000) of
    800 m x 800 m at dh = 20/s m with a 160 m C-PML (8 s cells), dt = 2/s ms,
    0.8 s; source at 100 m depth near the left edge, receivers along 60 m
    depth, so that forward and adjoint waves run through the C-PML; observed
    data from a +5% Gaussian Vp/Vs anomaly.  Returns the relative gap between
    the lambda gradient's derivative along that anomaly and the central
    difference of the misfit (FP64)."""
    dh, nPml, n = 20.0 / s, 8 * s, 40 * s
    dt, nt, f0 = 2e-3 / s, 400 * s, 8.0
    NZ = NX = n + 2 * nPml
    g = oracle_mod.geom(NZ, NX, nPml, 0, nt, dh, dh, dt, f0)
    zz, xx = np.mgrid[0:NZ, 0:NX]
    Z, X = (zz - nPml) * dh, (xx - nPml) * dh
    blob = np.exp(-((Z - 480.0) ** 2 + (X - 400.0) ** 2) / (2 * 80.0 ** 2))

    def moduli(vp):
        vs, rho = vp / np.sqrt(3.0), np.full((NZ, NX), 2000.0)
        return rho * (vp ** 2 - 2 * vs ** 2) / 1e6, rho * vs ** 2 / 1e6, rho

    vp = np.full((NZ, NX), 3000.0)
    src = [(nPml + int(round(100 / dh)), nPml + int(round(200 / dh)))]
    rx = nPml + np.arange(0, n, s)
    recs = [(np.full(len(rx), nPml + int(round(60 / dh))), rx)]
    stf = sourceGene(f0, nt, dt)
    lam, mu, rho = moduli(vp)
    h = 1e-4
    with oracle_mod.mutation(mutation):
        obs = oracle_mod.run(g, 2, *moduli(vp * (1 + 0.05 * blob)), stf, src, recs,
                             precision=64)["syn"]
        gL = oracle_mod.run(g, 1, lam, mu, rho, stf, src, recs, obs, precision=64)["grad_lambda"]
        J = [oracle_mod.run(g, 0, lam * (1 + sgn * h * blob), mu, rho, stf, src, recs, obs,
                            precision=64)["misfit"] for sgn in (1.0, -1.0)]
    fd = (J[0] - J[1]) / (2 * h)
    return abs(np.sum(gL * lam * blob) - fd) / abs(fd)


def test_cpml_adjoint_gap_shrinks_with_refinement(oracle_mod):
    """With waves in the C-PML, the reference's adjoint (el_velocity_adj.cu,
    el_stress_adj.cu) differs from the derivative of its forward operator:
    its C-PML coefficients sit outside the difference stencils, its memory
    variables are updated with the coefficients of the updated cell, and
    el_stress_adj's memory variables run ungated through the interior.  In a
    homogeneous model, refining the grid at fixed C-PML thickness (nPml
    doubled, dh and dt halved) cuts the gradient-vs-FD gap from 17.5% to
    5.3%; with el_velocity_adj's a_x * d(psi_xx)/dx term dropped the gap grows
    instead (37% -> 215%).  Half-cell shifts of the profiles (a_x_half for
    a_x, K_x_half for K_x) converge like the reference and are not visible
    to this check (measured)."""
    g1, g2 = _pml_fd_gap(oracle_mod, 1), _pml_fd_gap(oracle_mod, 2)
    assert g2 < 0.5 * g1 and g2 < 0.1
    m1 = _pml_fd_gap(oracle_mod, 1, "adj_vel_drop_psi")
    m2 = _pml_fd_gap(oracle_mod, 2, "adj_vel_drop_psi")
    assert m2 > m1 and m2 > 0.5


# ------------------------------------------------------------------ adjoint / gradient
def _deep_interior_case():
    nz, nx, nt, dh, dt, f0 = 120, 120, 150, 10.0, 1e-3, 25.0
    vp = np.full((nz, nx), 3000.0)
    cs = vp / np.sqrt(3)
    rho = 310 * vp ** 0.25
    return Case("deep", nz, nx, dh, dh, nt, dt, f0, 12, vp, cs, rho, vp.copy(), cs.copy(),
                rho.copy(), np.array([50, 60]), np.array([55, 65]), np.full(20, 70),
                np.arange(40, 80, 2), sourceGene(f0, nt, dt))


def test_gradient_matches_central_fd_where_adjoint_is_exact(oracle_mod, workspace):
    """In a homogeneous medium with all waves inside the interior the
    reference's hand-derived adjoint is the exact discrete adjoint, so the
    gradient equals the central finite difference of the misfit (FP64) for
    lambda, mu and rho: this pins time indexing, signs and the harmonic-mean /
    buoyancy chain rules (the atomic sprays)."""
    c = _deep_interior_case()
    para, stf = models.write_case(c, workspace)
    lam, mu, den = models.moduli(c, "true")
    rng = np.random.default_rng(1)
    NZ, NX = lam.shape
    v = np.zeros((NZ, NX))
    v[c.nPml + 20:NZ - c.nPad - c.nPml - 20, c.nPml + 20:NX - c.nPml - 20] = 1
    pert = gaussian_filter(rng.standard_normal((NZ, NX)), 4) * v
    pert /= np.abs(pert).max()
    ids = [0, 1]
    oracle_mod.cufd(lam * (1 + 0.05 * pert), mu * (1 + 0.05 * pert), den, stf, 2, ids, para,
                    precision=64)
    _, gL, gM, gD, _, _ = oracle_mod.cufd(lam, mu, den, stf, 1, ids, para, precision=64)
    J = lambda L_, M_, D_: oracle_mod.cufd(L_, M_, D_, stf, 0, ids, para, precision=64)[0]
    for k, (g, scale) in enumerate(((gL, 1.0), (gM, 1.0), (gD, 10.0))):
        for trial in range(2):
            d = gaussian_filter(rng.standard_normal((NZ, NX)), 3) * v
            d *= scale / np.abs(d).max()
            h = 1e-2
            p = [lam.copy(), mu.copy(), den.copy()]
            m = [lam.copy(), mu.copy(), den.copy()]
            p[k] += h * d
            m[k] -= h * d
            fd = (J(*p) - J(*m)) / (2 * h)
            assert np.sum(g * d) == pytest.approx(fd, rel=1e-7)


def test_taylor_remainder_is_second_order(oracle_mod, workspace):
    """deps/CustomOps/FWI/gradtest.jl:137-151: |f(m+gv)-f(m)| = O(g) and
    |f(m+gv)-f(m)-g<v,grad f>| = O(g^2), in the exact-adjoint setting."""
    c = _deep_interior_case()
    para, stf = models.write_case(c, workspace)
    lam, mu, den = models.moduli(c, "true")
    ids = [0, 1]
    oracle_mod.cufd(lam * 1.02, mu, den, stf, 2, ids, para, precision=64)
    _, gL, gM, gD, _, _ = oracle_mod.cufd(lam, mu, den, stf, 1, ids, para, precision=64)
    NZ, NX = lam.shape
    v = np.zeros((NZ, NX))
    v[c.nPml + 20:NZ - c.nPad - c.nPml - 20, c.nPml + 20:NX - c.nPml - 20] = 1
    d = gaussian_filter(np.random.default_rng(3).standard_normal((NZ, NX)), 3) * v
    d /= np.abs(d).max()
    f0 = oracle_mod.cufd(lam, mu, den, stf, 0, ids, para, precision=64)[0]
    gs = [10.0, 3.0, 1.0]
    s, w = [], []
    for gam in gs:
        f = oracle_mod.cufd(lam + gam * d, mu, den, stf, 0, ids, para, precision=64)[0]
        s.append(abs(f - f0))
        w.append(abs(f - f0 - gam * np.sum(gL * d)))
    rs = np.log(s[0] / s[-1]) / np.log(gs[0] / gs[-1])
    rw = np.log(w[0] / w[-1]) / np.log(gs[0] / gs[-1])
    assert rs == pytest.approx(1.0, abs=0.15)
    assert rw == pytest.approx(2.0, abs=0.15)


def test_matches_independent_numpy_restatement(oracle_mod, workspace):
    """Tiny heterogeneous grid with the C-PML, the un-gated adjoint memory
    variables and both sprays active: the C++ oracle and the vectorised numpy
    transcription agree in FP64."""
    c = models.case_small(nz=24, nx=30, nSteps=120, nshots=2, nPml=6)
    para, stf, ids = _setup(oracle_mod, c, workspace)
    lam0, mu0, den0 = models.moduli(c, "start")
    J = oracle_mod.cufd(lam0, mu0, den0, stf, 0, ids, para, precision=64)[0]
    _, gL, gM, gD, gS, r = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=64)
    op = NR.Op(c.NZ, c.NX, c.nPml, c.nPad, c.dz, c.dx, c.dt, c.f0, lam0, mu0, den0)
    Jn, G, S = 0.0, None, []
    for s in range(2):
        src = (c.src_z[s] + c.nPml, c.src_x[s] + c.nPml)
        recs = (c.rec_z + c.nPml, c.rec_x + c.nPml)
        obs = np.fromfile(os.path.join(workspace, "Data", f"Shot{s}.bin"), dtype=np.float32)
        obs = obs.reshape(len(c.rec_x), c.nSteps).astype(np.float64)
        syn, j, g = op.run(src, recs, stf[s], obs)
        np.testing.assert_allclose(r["syn"][s], syn, rtol=0, atol=1e-6 * np.abs(syn).max())
        Jn += j
        S.append(g["stf"])
        G = g if G is None else {k: G[k] + g[k] for k in ("lam", "mu", "den")}
    assert J == pytest.approx(Jn, rel=1e-12)
    for a, k in ((gL, "lam"), (gM, "mu"), (gD, "den")):
        assert np.linalg.norm(a - G[k]) <= 1e-12 * np.linalg.norm(a)
    np.testing.assert_allclose(gS[:2], np.stack(S), rtol=0, atol=1e-12 * np.abs(gS).max())


def test_illumination_matches_independent_numpy_restatement(oracle_mod, workspace):
    """Forward illumination (sum over shots and steps of vz^2 + vx^2 on the
    update region): the C++ oracle against the numpy transcription in FP64, and
    zero outside the update region."""
    c = models.case_small(nz=24, nx=30, nSteps=120, nshots=2, nPml=6)
    para, stf, ids = _setup(oracle_mod, c, workspace)
    lam0, mu0, den0 = models.moduli(c, "start")
    r = oracle_mod.cufd(lam0, mu0, den0, stf, 2, ids, para, precision=64, illumination=True)[5]
    op = NR.Op(c.NZ, c.NX, c.nPml, c.nPad, c.dz, c.dx, c.dt, c.f0, lam0, mu0, den0)
    il = np.zeros((c.NZ, c.NX))
    for s in range(2):
        op.run((c.src_z[s] + c.nPml, c.src_x[s] + c.nPml), (c.rec_z + c.nPml, c.rec_x + c.nPml),
               stf[s])
        il += op.illum
    assert np.linalg.norm(r["illum"] - il) <= 1e-12 * np.linalg.norm(il)
    assert np.all(r["illum"][~op.U] == 0.0) and np.all(r["illum"] >= 0.0)
    assert np.count_nonzero(r["illum"][op.U]) > 0.9 * np.count_nonzero(op.U)


def test_backward_schemes_agree(oracle_mod, workspace):
    """Boundary-ring time reversal (the reference) vs checkpoint recompute."""
    c = models.case_small(nSteps=300)
    para, stf, ids = _setup(oracle_mod, c, workspace)
    lam0, mu0, den0 = models.moduli(c, "start")
    a = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=64, scheme=0)
    b = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=64, scheme=1, ckpt_every=17)
    for x, y in zip(a[1:5], b[1:5]):
        assert np.linalg.norm(x - y) <= 1e-12 * np.linalg.norm(y)
    a32 = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=32, scheme=0)
    for x, y in zip(a32[1:4], b[1:4]):
        assert np.linalg.norm(x - y) <= 1e-4 * np.linalg.norm(y)


def test_interior_resets_restore_the_forward_states(oracle_mod, workspace):
    """Scheme 2 (ring reversal plus interior resets every k steps) against the
    exact checkpoint-recompute scheme in FP64: a reset that wrote back any
    other step's state would move the gradient far beyond round-off, and the
    distance to the exact scheme shrinks as resets get denser."""
    c = models.case_small(nSteps=300)
    para, stf, ids = _setup(oracle_mod, c, workspace)
    lam0, mu0, den0 = models.moduli(c, "start")
    ex = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=64, scheme=1,
                         ckpt_every=25)
    err = {}
    for k in (0, 1, 7, 60):
        r = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=64,
                            scheme=2 if k else 0, ckpt_every=k)
        err[k] = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(r[1:5], ex[1:5]))
        assert err[k] < 1e-13, (k, err)
    assert err[1] <= err[0]


def test_fp32_reference_arithmetic_close_to_fp64(oracle_mod, workspace):
    c = models.case_small(nSteps=300)
    para, stf, ids = _setup(oracle_mod, c, workspace)
    lam0, mu0, den0 = models.moduli(c, "start")
    r32 = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=32)
    r64 = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=64)
    for s in range(2):
        d32, d64 = r32[5]["syn"][s], r64[5]["syn"][s]
        assert np.linalg.norm(d32 - d64) <= 1e-5 * np.linalg.norm(d64)
    for x, y in zip(r32[1:4], r64[1:4]):
        assert np.linalg.norm(x - y) <= 1e-4 * np.linalg.norm(y)


def test_reference_adjoint_is_approximate_in_heterogeneous_media(oracle_mod, workspace):
    """Documents a property of the reference: its adjoint evaluates material
    parameters at the updated cell instead of inside the difference stencils
    and uses its own C-PML adjoint, so in a heterogeneous model next to the
    C-PML its gradient differs from the FD derivative by a few percent.  The
    GPU path reproduces the reference's gradient, not the exact one."""
    c = models.case_small(nz=40, nx=60, nSteps=400, nshots=2)
    para, stf, ids = _setup(oracle_mod, c, workspace)
    lam0, mu0, den0 = models.moduli(c, "start")
    _, gL, _, _, _, _ = oracle_mod.cufd(lam0, mu0, den0, stf, 1, ids, para, precision=64)
    NZ, NX = lam0.shape
    v = np.zeros((NZ, NX))
    v[c.nPml + 5:NZ - c.nPad - c.nPml - 5, c.nPml + 5:NX - c.nPml - 5] = 1
    d = gaussian_filter(np.random.default_rng(0).standard_normal((NZ, NX)), 3) * v
    d /= np.abs(d).max()
    h = 1e-3
    J = lambda L_: oracle_mod.cufd(L_, mu0, den0, stf, 0, ids, para, precision=64)[0]
    fd = (J(lam0 + h * d) - J(lam0 - h * d)) / (2 * h)
    rel = abs(np.sum(gL * d) - fd) / abs(fd)
    assert 1e-4 < rel < 0.2


def test_reference_adjoint_grows_with_a_zero_shear_layer(oracle_mod, workspace):
    """Property of the restated reference scheme (not of this engine): its
    hand-derived adjoint applies the material parameters at the updated cell
    instead of inside the difference stencils, which is not the transpose of
    the forward operator at a fluid-solid contrast.  With Vs = 0 in the top
    rows the adjoint field (and the gradient) grows exponentially with the
    record length, while the same model with an elastic top stays bounded.
    The engine reproduces this scheme for parity; inversion experiments use
    models without a Vs = 0 layer."""
    rms = {}
    for water in (True, False):
        for nt in (400, 1200):
            c = models.case_small(nz=40, nx=60, nSteps=nt, nshots=1)
            if water:
                for a in (c.cs, c.cs0):
                    a[:5] = 0.0
            ws = os.path.join(workspace, f"w{int(water)}_{nt}")
            para, stf = models.write_case(c, ws)
            lam, mu, den = models.moduli(c, "true")
            oracle_mod.cufd(lam, mu, den, stf, 2, [0], para, precision=64)
            l0, m0, d0 = models.moduli(c, "start")
            gL = oracle_mod.cufd(l0, m0, d0, stf, 1, [0], para, precision=64)[1]
            rms[(water, nt)] = float(np.sqrt(np.mean(gL ** 2)))
    assert rms[(True, 1200)] > 1e3 * rms[(True, 400)], rms
    assert rms[(False, 1200)] < 1e2 * rms[(False, 400)], rms


@pytest.mark.parametrize("water", [False, True], ids=["heterogeneous", "zero_shear_layer"])
def test_exact_adjoint_equals_central_differences(oracle_mod, workspace, water):
    """Scheme 3 (no reference counterpart): the transpose of the forward time
    step, on the forward states of the exact recompute scheme.  Heterogeneous
    model, waves in the C-PML, 400 steps; also with Vs = 0 in the top 5
    rows, where the reference's adjoint is unstable: the lambda, mu, rho and
    source-function gradients equal central differences of the FP64 misfit
    to 1e-7 in smooth random directions (the reference's scheme on the same
    directions: 4%-15%, and its stf gradient has neither the 1500^2 scale nor
    the source window's chain rule, utilities.cu:582-593)."""
    c = models.case_small(nz=40, nx=60, nSteps=400, nshots=2)
    if water:
        for a in (c.cs, c.cs0):
            a[:5] = 0.0
    para, stf, ids = _setup(oracle_mod, c, workspace)
    m0 = list(models.moduli(c, "start"))
    NZ, NX = m0[0].shape
    v = np.zeros((NZ, NX))
    top = c.nPml + (7 if water else 4)  # mu = 0 rows are not differentiable in mu
    v[top:NZ - c.nPad - c.nPml - 4, c.nPml + 4:NX - c.nPml - 4] = 1
    rng = np.random.default_rng(5)
    _, gL, gM, gD, gS, _ = oracle_mod.cufd(*m0, stf, 1, ids, para, precision=64, scheme=3,
                                           ckpt_every=50)
    J = lambda *m: oracle_mod.cufd(*m, stf, 0, ids, para, precision=64)[0]  # noqa: E731
    h = 1e-3
    for k, (g, sc) in enumerate(((gL, 1.0), (gM, 1.0), (gD, 10.0))):
        d = gaussian_filter(rng.standard_normal((NZ, NX)), 3) * v
        d *= sc / np.abs(d).max()
        mp = [a + h * d if q == k else a for q, a in enumerate(m0)]
        mm = [a - h * d if q == k else a for q, a in enumerate(m0)]
        fd = (J(*mp) - J(*mm)) / (2 * h)
        assert np.sum(g * d) == pytest.approx(fd, rel=1e-7), k
    ds = np.zeros_like(stf)
    ds[:, 20:-20] = rng.standard_normal((stf.shape[0], stf.shape[1] - 40))
    Js = lambda s: oracle_mod.cufd(*m0, s, 0, ids, para, precision=64)[0]  # noqa: E731
    fd = (Js(stf + h * ds) - Js(stf - h * ds)) / (2 * h)
    assert np.sum(gS * ds) == pytest.approx(fd, rel=1e-7)
