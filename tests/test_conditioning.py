"""The oracle's data-conditioning stages (the optional residual stages of
libCUFD.cu:254-309), pinned against a library FFT, closed forms and the
identities each stage must satisfy.  CPU only.

Each check is chosen so that a plausible slip in the restatement fails it:
the frequency step (1/(2 nt dt), not 1/(nt dt)), amp^2 rather than amp, the
1/(2 nt) normalisation, the skipped Nyquist bin, the window ramps and clamps,
the source-update ratio and the adjoint's use of coef rather than conj(coef).
"""
import numpy as np
import pytest

from oracle import oracle

DT = 0.002  # 500 Hz sampling, Nyquist 250 Hz


def _tone(nt, f, dt=DT):
    t = np.arange(nt) * dt
    return np.hanning(nt) * np.sin(2 * np.pi * f * t)


@pytest.mark.parametrize("n", [1, 2, 7, 12, 30, 64, 100, 126, 2002, 2018, 8000])
def test_dft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    tol = 1e-11 * max(1.0, np.abs(ref).max())
    assert np.abs(oracle.dft(x, -1) - ref).max() < tol
    assert np.abs(oracle.dft(x, +1) - np.fft.ifft(x) * n).max() < tol


def test_bandpass_all_pass_is_identity():
    """A response of 1 on every bin: pad, transform, back, crop, 1/(2 nt)."""
    x = np.random.default_rng(0).standard_normal((3, 250))
    for prec in (32, 64):
        y = oracle.bandpass(x, DT, [-2.0, -1.0, 1e6, 2e6], precision=prec)
        assert np.abs(y - x).max() < (1e-5 if prec == 32 else 1e-12)


def test_bandpass_pass_stop_and_transition():
    nt = 1000
    f = [5.0, 10.0, 40.0, 60.0]
    for prec in (32, 64):
        # passband: unchanged; stopband: removed
        x = _tone(nt, 20.0)
        y = oracle.bandpass(x, DT, f, precision=prec)[0]
        assert np.linalg.norm(y - x) < 2e-3 * np.linalg.norm(x)
        x = _tone(nt, 100.0)
        y = oracle.bandpass(x, DT, f, precision=prec)[0]
        assert np.linalg.norm(y) < 2e-3 * np.linalg.norm(x)
        # mid-taper (50 Hz): amp = cos(pi/4), the data are scaled by amp^2 = 0.5
        x = _tone(nt, 50.0)
        y = oracle.bandpass(x, DT, f, precision=prec)[0]
        ratio = np.dot(y, x) / np.dot(x, x)
        assert abs(ratio - 0.5) < 0.02, ratio


def test_bandpass_nyquist_bin_skipped_when_nt_is_a_multiple_of_32():
    """The reference's frequency kernel runs on nt columns (of 32 threads):
    bin nt (Nyquist of the 2 nt grid) keeps its value when 32 | nt.  An
    alternating trace puts X[nt] = nt there, i.e. 0.5 (-1)^t after 1/(2 nt)."""
    f = [5.0, 10.0, 40.0, 60.0]
    for nt, want in ((64, 0.5), (70, 0.0), (96, 0.5), (100, 0.0)):
        a = (-1.0) ** np.arange(nt)
        y = oracle.bandpass(a, DT, f, precision=64)[0]
        proj = np.dot(y, a) / nt
        assert abs(proj - want) < 0.05, (nt, proj)


def test_window_traces_closed_form():
    nt = 400
    x = np.ones((4, nt))
    ws = np.array([0.1, 0.1, 0.3, -1.0])
    we = np.array([0.5, 0.5, 0.3, 100.0])
    wt = np.array([1.0, 2.5, 1.0, 1.0])
    for prec in (32, 64):
        y = oracle.window_traces(x, DT, ws, we, wt, 0.1, precision=prec)
        # trace 0: ramps [0.1, 0.14) and [0.46, 0.5): mid-ramp sin^2(pi/4) = 0.5
        t = np.arange(nt) * DT
        assert np.all(y[0][t < 0.099] == 0)
        assert abs(y[0][60] - 0.5) < 1e-3      # t = 0.12
        assert y[0][150] == pytest.approx(1.0)  # t = 0.30
        assert abs(y[0][240] - 0.5) < 1e-3     # t = 0.48
        assert np.all(y[0][t > 0.501] == 0)
        # weights multiply
        assert np.allclose(y[1], 2.5 * y[0], rtol=1e-6)
        # an empty window leaves the trace as it is (the kernel returns early)
        assert np.all(y[2] == 1.0)
        # clamped to [0, nt dt]: the default window with the same ratio
        ref = oracle.window(x[3:], DT, 0.1, precision=prec)
        assert np.abs(y[3] - ref[0]).max() < 1e-6
    # ratio > 0.5: every element left as it is
    y = oracle.window_traces(x, DT, ws, we, wt, 0.6)
    assert np.all(y == 1.0)


def _pulses(nrec, nt, seed, lo, hi):
    """Smooth random traces supported in samples [lo, hi)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((nrec, nt))
    k = np.hanning(15)
    for r in range(nrec):
        v = np.zeros(nt)
        v[lo:hi] = np.convolve(rng.standard_normal(hi - lo), k, mode="same") * np.hanning(hi - lo)
        out[r] = v
    return out


def test_source_update_recovers_scale_and_delay():
    nt, nrec = 300, 6
    cal = _pulses(nrec, nt, 1, 40, 180)  # zero in the 0.01-ratio ramp (first 6 samples)
    src = _pulses(1, nt, 2, 20, 120)[0]
    # obs = 1.7 cal: coef = 1.7 on every bin that carries energy
    cal2, src2, coef = oracle.source_update(1.7 * cal, cal, src, DT, precision=64)
    # (bins with energy near the 1e-6 regulariser are shrunk towards 0)
    assert np.linalg.norm(cal2 - 1.7 * cal) < 1e-4 * np.linalg.norm(cal)
    assert np.linalg.norm(src2 - 1.7 * src) < 1e-3 * np.linalg.norm(src)
    # obs = cal delayed by 7 samples: coef = exp(-2 pi i k 7 / (2 nt)), the
    # updated source is the source delayed by 7 samples
    d = 7
    obs = np.zeros_like(cal)
    obs[:, d:] = cal[:, :-d]
    cal2, src2, coef = oracle.source_update(obs, cal, src, DT, precision=64)
    assert np.linalg.norm(cal2 - obs) < 1e-4 * np.linalg.norm(obs)
    want = np.zeros_like(src)
    want[d:] = src[:-d]
    assert np.linalg.norm(src2 - want) < 1e-3 * np.linalg.norm(src)
    # the adjoint application multiplies by coef itself (utilities.cu:1068),
    # not conj(coef): an impulse is delayed by 7 samples as well
    imp = np.zeros((1, nt))
    imp[0, 100] = 1.0
    out = oracle.source_update_adj(imp, coef, DT, precision=64)[0]
    assert np.argmax(np.abs(out)) == 100 + d
    assert abs(out[100 + d] - 1.0) < 0.05
    # FP32 restatement agrees with FP64
    c32, s32, _ = oracle.source_update(obs, cal, src, DT, precision=32)
    assert np.linalg.norm(c32 - cal2) < 1e-5 * np.linalg.norm(cal2)


def _cond_case(nSteps=300):
    from fwiflow import models
    c = models.case_small(nz=30, nx=44, nSteps=nSteps, nshots=2, nPml=8)
    g = oracle.geom(c.NZ, c.NX, c.nPml, c.nPad, c.nSteps, c.dz, c.dx, c.dt, c.f0)
    lam, mu, den = models.moduli(c, "true")
    lam0, mu0, den0 = models.moduli(c, "start")
    src = [(z + c.nPml, x + c.nPml) for z, x in zip(c.src_z, c.src_x)]
    recs = [(c.rec_z + c.nPml, c.rec_x + c.nPml)] * len(src)
    stf = np.repeat(c.stf.reshape(1, -1), len(src), axis=0)
    nrec = len(c.rec_x)
    T = c.nSteps * c.dt
    cond = {"if_win": True, "filter": [2.0, 5.0, 25.0, 40.0], "if_src_update": True,
            "win_start": [0.02 + 0.001 * np.arange(nrec) + 0.01 * s for s in range(len(src))],
            "win_end": [np.full(nrec, 0.85 * T) - 0.002 * s for s in range(len(src))],
            "weights": [1.0 + 0.01 * np.arange(nrec) + 0.1 * s for s in range(len(src))]}
    return c, g, (lam, mu, den), (lam0, mu0, den0), src, recs, stf, cond


def test_operator_applies_the_stages_in_the_reference_order():
    """libCUFD.cu:254-309: window obs and syn, band-pass both, source update,
    res = obs - syn (sample 0 zeroed), misfit, then on res: source-update
    adjoint, band-pass, window with ratio 0.1.  The operator's outputs equal
    that composition of the separately pinned stages."""
    c, g, true, start, src, recs, stf, cond = _cond_case()
    obs = oracle.run(g, 2, *true, stf, src, recs, precision=64)["syn"]
    r = oracle.run(g, 0, *start, stf, src, recs, obs, precision=64, cond=cond)
    dt = c.dt
    J = 0.0
    for s in range(len(src)):
        ws, we, wt = cond["win_start"][s], cond["win_end"][s], cond["weights"][s]
        o = oracle.window_traces(obs[s], dt, ws, we, wt, 0.005, precision=64)
        y = oracle.window_traces(r["syn"][s], dt, ws, we, wt, 0.005, precision=64)
        o = oracle.bandpass(o, dt, cond["filter"], precision=64)
        y = oracle.bandpass(y, dt, cond["filter"], precision=64)
        srow = oracle.window(stf[s], dt, 0.001, precision=64)[0]
        y, srow, coef = oracle.source_update(o, y, srow, dt, precision=64)
        res = o - y
        res[:, 0] = 0.0
        J += 0.5 * np.sum(res * res)
        res = oracle.source_update_adj(res, coef, dt, precision=64)
        res = oracle.bandpass(res, dt, cond["filter"], precision=64)
        res = oracle.window_traces(res, dt, ws, we, wt, 0.1, precision=64)
        scale = np.abs(o).max()
        assert np.abs(r["cond_obs"][s] - o).max() < 1e-6 * scale
        assert np.abs(r["cond_syn"][s] - y).max() < 1e-6 * scale
        assert np.abs(r["src_updated"][s] - srow).max() < 1e-6 * np.abs(srow).max()
        assert np.abs(r["res"][s] - res).max() < 1e-6 * np.abs(res).max()
    assert r["misfit"] == pytest.approx(J, rel=1e-6)  # syn is returned rounded to f32


def test_operator_fp32_close_to_fp64_with_conditioning():
    c, g, true, start, src, recs, stf, cond = _cond_case()
    obs = oracle.run(g, 2, *true, stf, src, recs, precision=32)["syn"]
    r64 = oracle.run(g, 1, *start, stf, src, recs, obs, precision=64, cond=cond)
    r32 = oracle.run(g, 1, *start, stf, src, recs, obs, precision=32, cond=cond)
    for k in ("grad_lambda", "grad_mu", "grad_den"):
        a, b = r32[k], r64[k]
        assert np.linalg.norm(a - b) < 1e-3 * np.linalg.norm(b), k
    for s in range(len(src)):
        a, b = r32["res"][s], r64["res"][s]
        assert np.linalg.norm(a - b) < 1e-4 * np.linalg.norm(b)
