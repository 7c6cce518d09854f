"""GPU parity of the sliding-window bleach mode (IFCS:9146-9172) against the
oracle's restatement of correlate()'s window loop: averaged ACF within 1e-4
relative (floor 1e-4 of the amplitude), averaged blocked variance within
1e-4 relative for pixels whose every window blocks alike (at least 95 % of
them), fitted N and D within 1e-3 relative where both sides fitted."""
import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
import oracle
from parity import rel_err

pytestmark = pytest.mark.gpu

PIX, PSF = 0.24e-6, 0.38e-6


@pytest.mark.parametrize("L", [2000, 1000])   # correlator depth 7 (main kernel path) and 6 (generic path)
def test_sliding_window_matches_oracle(engine, L):
    stack = imfcs.simulate_2d(8, 6, 12000, 50, D_um2s=1.0, tau_bleach_s=10.0, seed=1234 + L)
    T, H, W = stack.shape
    bg = int(stack.min())
    acf_m = imfcs.make_fit_model(PIX, PIX, PSF)
    cfg = imfcs.make_pipeline_config(W, H, T, binning=1, background=bg, poly_order=0, do_fit=True,
                                     acf_model=acf_m, sliding_window=L)
    engine.upload(stack)
    engine.run(cfg)
    g = engine.results(cfg)
    q = oracle.sliding_window_q(L)
    n = 16 + (q - 1) * 8 + 1
    assert g["acf"].shape[-1] == n
    _, _, _, lagtime = oracle.tables(16, q, L, 1e-3)
    om = oracle.make_model(PIX, PIX, PSF)
    o_acf, o_var = np.zeros((H, W, n)), np.zeros((H, W, n))
    o_fit, o_ok = np.full((H, W, 11), np.nan), np.zeros((H, W), bool)
    for y in range(H):
        for x in range(W):
            a, v, _, _, _, qq = oracle.sliding_window_cf(stack, x, y, 1, 1, bg, 1, T, L)
            assert qq == q
            o_acf[y, x], o_var[y, x] = a, v
            fr, _, ok, _, _ = oracle.standard_fit(om, lagtime, a, v)
            o_fit[y, x], o_ok[y, x] = fr, ok
    ga, gv = g["acf"][0], g["var"][0]
    amp = np.abs(o_acf[..., 1:]).max(-1, keepdims=True)
    e = rel_err(ga, o_acf, 1e-4 * amp)
    assert e.max() < 1e-4, e.max()
    ev = rel_err(gv[..., 1:], o_var[..., 1:], 1e-300).max(-1)
    assert (ev < 1e-4).mean() >= 0.95, np.sort(ev)[-5:]
    ok = (g["states"][0] == 0) & o_ok
    assert ok.mean() > 0.8
    for k in (0, 1):
        err = rel_err(g["fitres"][0][..., k][ok], o_fit[..., k][ok], 1e-30)
        assert err.max() < 1e-3, (k, err.max())


def test_linear_segments_match_oracle(engine):
    """"Lin Segment" correction (IFCS:9767-9805) on the GPU against the
    oracle: segment means and bcInt0 to 1e-12 relative, ACF within 1e-4,
    blocked variance within 1e-4 where the blocking index agrees, fitted N
    and D within 1e-3."""
    stack = imfcs.simulate_2d(8, 6, 12000, 60, D_um2s=1.0, tau_bleach_s=8.0, seed=4321)
    T, H, W = stack.shape
    bg = int(stack.min())
    L = 600
    acf_m = imfcs.make_fit_model(PIX, PIX, PSF)
    cfg = imfcs.make_pipeline_config(W, H, T, binning=1, background=bg, poly_order=0, do_fit=True,
                                     acf_model=acf_m, lin_segment=L)
    engine.upload(stack)
    engine.run(cfg)
    g = engine.results(cfg)
    _, _, _, lagtime = oracle.tables(16, 8, T, 1e-3)
    om = oracle.make_model(PIX, PIX, PSF)
    n = lagtime.size
    o_acf, o_var = np.zeros((H, W, n)), np.zeros((H, W, n))
    o_ind = np.zeros((H, W), int)
    o_fit, o_ok = np.full((H, W, 11), np.nan), np.zeros((H, W), bool)
    for y in range(H):
        for x in range(W):
            d = oracle.get_intensity(stack, x, y, 1, 1, bg, 1, T)
            d, seg = oracle.bleach_linseg(d, L)
            np.testing.assert_allclose(g["coef"][y * W + x], seg, rtol=1e-12)
            ind, _ = oracle.block_transform(d, d)
            a, v, _, _ = oracle.calculate_cf(d, d, ind)
            o_acf[y, x], o_var[y, x], o_ind[y, x] = a, v, ind
            fr, _, ok, _, _ = oracle.standard_fit(om, lagtime, a, v)
            o_fit[y, x], o_ok[y, x] = fr, ok
    # the reference's extrapolation after the last segment centre can turn
    # negative under strong bleaching: those pixels are NaN on both sides
    fin = np.isfinite(o_acf).all(-1)
    assert np.array_equal(np.isfinite(g["acf"][0]).all(-1), fin)
    assert fin.mean() > 0.5
    amp = np.abs(o_acf[..., 1:]).max(-1, keepdims=True)
    e = rel_err(g["acf"][0][fin], o_acf[fin], 1e-4 * amp[fin])
    assert e.max() < 1e-4, e.max()
    same = (g["ind"][0] == o_ind) & fin
    assert same[fin].mean() >= 0.95
    ev = rel_err(g["var"][0][same][:, 1:], o_var[same][:, 1:], 1e-300)
    assert ev.max() < 1e-4, ev.max()
    ok = (g["states"][0] == 0) & o_ok & same
    assert ok[fin].mean() > 0.8
    for k in (0, 1):
        err = rel_err(g["fitres"][0][..., k][ok], o_fit[..., k][ok], 1e-30)
        assert err.max() < 1e-3, (k, err.max())


def test_single_exponential_correction_matches_oracle(engine):
    """"Single Exp" correction (IFCS:9682-9709): the GPU's FP64 LM fit of
    every trace against the oracle's, compared through the fitted curve at
    the trace points (1e-5 relative), then the ACF within 1e-4 relative
    (floor 1e-2 of the amplitude) and the fitted N, D within 1e-3.  Traces whose least-squares exponential degenerates
    into a line (t1 beyond 100 trace lengths, A and -C growing together) have
    no finite optimum: every LM stops somewhere along that valley, so they
    are left out of the comparison."""
    stack = imfcs.simulate_2d(8, 6, 20000, 70, D_um2s=1.0, tau_bleach_s=12.0, seed=78)
    T, H, W = stack.shape
    bg = int(stack.min())
    acf_m = imfcs.make_fit_model(PIX, PIX, PSF)
    cfg = imfcs.make_pipeline_config(W, H, T, binning=1, background=bg, poly_order=0, do_fit=True,
                                     acf_model=acf_m, exp_bleach=True)
    engine.upload(stack)
    engine.run(cfg)
    g = engine.results(cfg)
    _, _, _, lagtime = oracle.tables(16, 8, T, 1e-3)
    om = oracle.make_model(PIX, PIX, PSF)
    n = lagtime.size
    o_acf = np.zeros((H, W, n))
    o_ind = np.zeros((H, W), int)
    o_fit, o_ok = np.full((H, W, 11), np.nan), np.zeros((H, W), bool)
    posed = np.zeros((H, W), bool)
    curve = lambda p, t: p[0] * np.exp(-t / p[1]) + p[2]
    for y in range(H):
        for x in range(W):
            tr, it = oracle.intensity_trace(stack, x, y, 1, 1, bg, 1, T, 1000, 1e-3)
            res, _ = oracle.expfit(tr, it, 1)
            posed[y, x] = 0 < res[1] < 100 * T * 1e-3
            if posed[y, x]:
                gp = g["coef"][y * W + x][1:]
                np.testing.assert_allclose(curve(gp, it), curve(res, it), rtol=1e-5)
            d = oracle.get_intensity(stack, x, y, 1, 1, bg, 1, T)
            d = oracle.bleach_exp(d, res, 1, 1e-3)
            ind, _ = oracle.block_transform(d, d)
            a, v, _, _ = oracle.calculate_cf(d, d, ind)
            o_acf[y, x], o_ind[y, x] = a, ind
            fr, _, ok, _, _ = oracle.standard_fit(om, lagtime, a, v)
            o_fit[y, x], o_ok[y, x] = fr, ok
    assert posed.mean() > 0.8
    fin = np.isfinite(o_acf).all(-1) & posed
    assert np.array_equal(np.isfinite(g["acf"][0]).all(-1)[posed], np.isfinite(o_acf).all(-1)[posed])
    amp = np.abs(o_acf[..., 1:]).max(-1, keepdims=True)
    # floor 1e-2 of the amplitude: the two fits agree to ~1e-6 in their
    # parameters (nonlinear least squares, each stopped by its own test),
    # which moves channels near zero by ~1e-7 of the amplitude
    e = rel_err(g["acf"][0][fin], o_acf[fin], 1e-2 * amp[fin])
    print("ACF: max relative error %.2e (floor 1e-2 of the amplitude), %.2e (floor 1e-4)"
          % (e.max(), rel_err(g["acf"][0][fin], o_acf[fin], 1e-4 * amp[fin]).max()))
    assert e.max() < 1e-4, e.max()
    same = (g["ind"][0] == o_ind) & fin
    ok = (g["states"][0] == 0) & o_ok & same
    assert ok[fin].mean() > 0.8
    for k in (0, 1):
        err = rel_err(g["fitres"][0][..., k][ok], o_fit[..., k][ok], 1e-30)
        assert err.max() < 1e-3, (k, err.max())


def test_double_exponential_correction_matches_oracle(engine):
    """"Double Exp" correction (IFCS:9710-9738): a diffusing sample plus a
    deterministic two-exponential bleaching background (150 e^{-t/1 s} +
    150 e^{-t/6 s} photons per frame, Poisson), so that every trace has one
    well separated least-squares optimum.  The GPU's FP64 LM fit against
    the oracle's DoubleExpFit from the same initial guesses, compared
    through the fitted curve at the trace points (1e-5 relative); then the
    ACF within 1e-4 relative (floor 1e-2 of the amplitude, as for the single
    exponential) and the fitted N, D within 1e-3."""
    base = imfcs.simulate_2d(8, 6, 20000, 70, D_um2s=1.0, tau_bleach_s=0.0, counts_per_frame=50, seed=501)
    T, H, W = base.shape
    t = (np.arange(T) + 0.5) * 1e-3
    lam = 150 * np.exp(-t / 1.0) + 150 * np.exp(-t / 6.0)
    rng = np.random.default_rng(77)
    stack = (base.astype(np.int64) + rng.poisson(lam[:, None, None] * np.ones((1, H, W)))).astype(np.uint16)
    bg = int(stack.min())
    acf_m = imfcs.make_fit_model(PIX, PIX, PSF)
    cfg = imfcs.make_pipeline_config(W, H, T, binning=1, background=bg, poly_order=0, do_fit=True,
                                     acf_model=acf_m, exp_bleach=2)
    engine.upload(stack)
    engine.run(cfg)
    g = engine.results(cfg)
    assert g["coef"].shape == (H * W, 6)
    _, _, _, lagtime = oracle.tables(16, 8, T, 1e-3)
    om = oracle.make_model(PIX, PIX, PSF)
    n = lagtime.size
    o_acf = np.zeros((H, W, n))
    o_ind = np.zeros((H, W), int)
    o_fit, o_ok = np.full((H, W, 11), np.nan), np.zeros((H, W), bool)
    curve = lambda p, t: p[0] * np.exp(-t / p[1]) + p[2] * np.exp(-t / p[3]) + p[4]
    worst = 0.0
    for y in range(H):
        for x in range(W):
            tr, it = oracle.intensity_trace(stack, x, y, 1, 1, bg, 1, T, 1000, 1e-3)
            res, _ = oracle.expfit(tr, it, 2)
            assert 0.5 < min(res[1], res[3]) and max(res[1], res[3]) < 20, res
            gp = g["coef"][y * W + x]
            assert gp[0] == pytest.approx(gp[1] + gp[3] + gp[5], rel=1e-12)
            e = np.abs(curve(gp[1:], it) / curve(res, it) - 1).max()
            worst = max(worst, e)
            d = oracle.get_intensity(stack, x, y, 1, 1, bg, 1, T)
            d = oracle.bleach_exp(d, res, 2, 1e-3)
            ind, _ = oracle.block_transform(d, d)
            a, v, _, _ = oracle.calculate_cf(d, d, ind)
            o_acf[y, x], o_ind[y, x] = a, ind
            fr, _, ok, _, _ = oracle.standard_fit(om, lagtime, a, v)
            o_fit[y, x], o_ok[y, x] = fr, ok
    print("fitted bleach curves: max relative difference %.2e" % worst)
    assert worst < 1e-5
    fin = np.isfinite(o_acf).all(-1)
    assert fin.all()
    assert np.isfinite(g["acf"][0]).all()
    amp = np.abs(o_acf[..., 1:]).max(-1, keepdims=True)
    e = rel_err(g["acf"][0], o_acf, 1e-2 * amp)
    print("ACF: max relative error %.2e (floor 1e-2 of the amplitude), %.2e (floor 1e-4)"
          % (e.max(), rel_err(g["acf"][0], o_acf, 1e-4 * amp).max()))
    assert e.max() < 1e-4, e.max()
    same = g["ind"][0] == o_ind
    ok = (g["states"][0] == 0) & o_ok & same
    assert ok.mean() > 0.8
    for k in (0, 1):
        err = rel_err(g["fitres"][0][..., k][ok], o_fit[..., k][ok], 1e-30)
        assert err.max() < 1e-3, (k, err.max())
