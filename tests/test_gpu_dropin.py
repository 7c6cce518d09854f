"""GPU parity of the drop-in layer: the GpufitImFCS natives driven the way
Imaging_FCS_1_52.java drives them (through a fake JNIEnv where the JNI
boundary itself is under test, through the C ABI otherwise).

References: numpy for the integer/byte work (bit-exact), np.polyfit for the
linear least-squares bleach fit, closed-form model curves for the ACF fit,
and the oracle (the Java CPU path restated in C) for calcACF.
Tolerances: binning and trace averages bit-exact; polynomial coefficients
within 1e-4 relative (FP32 Gpufit semantics); ACF channels within 1e-4
relative (floor 1e-4 of the curve amplitude); fitted N, D within 1e-3.
"""
import ctypes

import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
from imaging_fcs_1_52_amd import plugin_glue
from fake_jni import ACF_FIELDS, FakeEnv
from parity import rel_err, run_oracle

pytestmark = pytest.mark.gpu

PIX, PSF = 0.24e-6, 0.38e-6


def _java_params(env, a):
    """ACFParameters ctypes mirror -> fake Java object with the same fields."""
    fields = {}
    for name, sig in ACF_FIELDS.items():
        v = getattr(a, name)
        fields[name] = bool(v) if sig == "Z" else (float(v) if sig == "D" else int(v))
    return env.acf_parameters(**fields)


def _params(W, H, T, binning, cf, poly_order, bg=0, frametime=1e-3, p=16, q=8):
    return plugin_glue.initialize_gpu_params(W, H, W - cf[0], H - cf[1], binning, cf, frametime, 1, T, bg,
                                             poly_order, p=p, q=q)


def test_calc_binning_is_exact_through_jni():
    rng = np.random.default_rng(3)
    W, H, T, b = 11, 9, 96, 3
    a = _params(W, H, T, b, (0, 0), 0)
    stack = rng.integers(0, 4096, size=(T, a.hin_star, a.win_star)).astype(np.float32)
    env = FakeEnv()
    out = np.zeros(a.w_temp * a.h_temp * T, np.float32)
    env.call("calcBinning", env.handle(stack.reshape(-1)), env.handle(out), _java_params(env, a))
    assert not env.exceptions and not env.trapped
    ref = np.zeros((T, a.h_temp, a.w_temp), np.float64)
    for k in range(b):
        for i in range(b):
            ref += stack[:, k:k + a.h_temp, i:i + a.w_temp]
    np.testing.assert_array_equal(out.reshape(T, a.h_temp, a.w_temp), ref.astype(np.float32))


def test_calc_data_bleach_correction_averages_through_jni():
    rng = np.random.default_rng(4)
    W, H, T = 6, 5, 5000
    a = _params(W, H, T, 1, (0, 0), 4)
    pix = rng.integers(0, 300, size=(T, a.h_temp, a.w_temp)).astype(np.float32)
    env = FakeEnv()
    out = np.zeros(a.w_temp * a.h_temp * a.nopit, np.float32)
    env.call("calcDataBleachCorrection", env.handle(pix.reshape(-1)), env.handle(out), _java_params(env, a))
    assert not env.exceptions and not env.trapped
    # calc_data_bleach_correction (agpufitjni.cu:4535-4553): double sum of
    # `ave` frames divided by ave, stored [y][x][nopit]
    means = pix[:a.nopit * a.ave].astype(np.float64).reshape(a.nopit, a.ave, a.h_temp, a.w_temp).sum(1) / a.ave
    np.testing.assert_array_equal(out.reshape(a.h_temp, a.w_temp, a.nopit),
                                  np.transpose(means, (1, 2, 0)).astype(np.float32))


def test_fit_linear_1d_matches_least_squares_through_jni():
    rng = np.random.default_rng(5)
    nfits, npts, ncoef = 64, 1000, 5
    # bleaching traces: exponential decays (tau 10-40 s) over 50 s plus noise
    t = (1e-3 * (np.arange(npts) + 0.5) * 50).astype(np.float32)
    amp = rng.uniform(80, 200, nfits)
    tau = rng.uniform(10, 40, nfits)
    data = amp[:, None] * np.exp(-t.astype(np.float64)[None, :] / tau[:, None]) + 20.0
    data = (data + rng.normal(0, 1.0, data.shape)).astype(np.float32)
    init = np.zeros((nfits, 11), np.float32)
    init[:, 0] = data[:, -1]
    pfit = np.array([1] * ncoef + [0] * (11 - ncoef), np.int32)
    out_p = np.zeros((nfits, 11), np.float32)
    out_s = np.zeros(nfits, np.int32)
    out_c = np.zeros(nfits, np.float32)
    out_i = np.zeros(nfits, np.int32)
    weights = np.ones_like(data)
    env = FakeEnv()
    h = [env.handle(x) for x in (data, weights, init, pfit, t, out_p, out_s, out_c, out_i)]
    st = env.call("fit", nfits, npts, h[0], h[1], imfcs.MODEL_LINEAR_1D, h[2], 1e-16, 2000, ncoef, h[3], 0,
                  npts * 4, h[4], h[5], h[6], h[7], h[8], restype=ctypes.c_int32)
    assert st == 0 and not env.trapped
    assert (out_s == 0).mean() > 0.9
    for k in range(nfits):
        if out_s[k] != 0:
            continue
        ref = np.polyfit(t.astype(np.float64), data[k].astype(np.float64), ncoef - 1)[::-1]
        got = out_p[k, :ncoef].astype(np.float64)
        # compare the fitted curves (the coefficients of a degree-4 fit are
        # strongly correlated, the curve is what the bleach correction uses)
        cg = np.polyval(got[::-1], t.astype(np.float64))
        cr = np.polyval(ref[::-1], t.astype(np.float64))
        assert np.max(np.abs(cg - cr)) / np.max(np.abs(cr)) < 1e-4


def test_fit_acf_1d_recovers_noise_free_model():
    import oracle
    om = oracle.make_model(PIX, PIX, PSF)
    lag, samp, mtab, lagtime = imfcs.correlator_tables(16, 8, 50000, 1e-3)
    x = lagtime[1:]
    truths = [(2.0, 1e-12, 0.001), (5.0, 3e-12, -0.002), (0.8, 0.5e-12, 0.0)]
    nfits, npts = len(truths), len(x)
    m = imfcs.FitModel(nfits, npts, True, imfcs.MODEL_ACF_1D, tolerance=np.float32(1e-16),
                       max_number_iterations=2000, num_valid_coefs=0,
                       parameters_to_fit=[True, True, False, False, True] + [False] * 15, user_info_size=npts * 4)
    geom = imfcs.make_fit_model(PIX, PIX, PSF)
    init = []
    for k, (N, D, G) in enumerate(truths):
        pareq = np.zeros(11)
        pareq[0], pareq[1], pareq[4], pareq[10] = N, D, G, 1e-5
        y = np.array([oracle.fcs_value(om, xi, pareq) for xi in x])
        m.data[k * npts:(k + 1) * npts] = y.astype(np.float32)
        init += [1.0, 1e-12, 0, 0, 0, 0, 0, 0, 0, 0, 1e-5, geom.ax, geom.ay, geom.s, geom.sz, geom.rx, geom.ry,
                 geom.fitobsvol, geom.q2, geom.q3]
    m.weights[:] = 1.0
    m.initial_parameters[:] = np.asarray(init, np.float32)
    m.user_info[:npts] = x.astype(np.float32)
    r = imfcs.fit(m)
    p = r.parameters.reshape(nfits, 20)
    for k, (N, D, G) in enumerate(truths):
        assert r.states[k] == 0
        assert abs(p[k, 0] - N) / N < 1e-3
        assert abs(p[k, 1] - D) / D < 1e-3
        assert abs(p[k, 4] - G) < 1e-3 / N


def _oracle_series(stack, a, px, py, bg, coef=None, frametime=1e-3):
    """1-based FP64 series of the window at raw (px, py), optionally bleach
    corrected at the reference GPU kernel's time origin t = frametime
    (i + 0.5), i = 0.. (agpufitjni.cu:5108)."""
    import oracle
    T = stack.shape[0]
    d = oracle.get_intensity(stack, px, py, a.binningX, a.binningY, bg, 1, T)
    if coef is not None:
        t = frametime * (np.arange(T) + 0.5)
        f = np.polynomial.polynomial.polyval(t, coef)
        r = np.sqrt(f / coef[0])
        d[1:] = d[1:] / r + coef[0] * (1 - r)
    return d


@pytest.mark.parametrize("poly_order,p,q", [(0, 16, 8), (4, 16, 8), (4, 32, 6)])
def test_calc_acf_through_jni_matches_oracle(poly_order, p, q):
    """calcACF with the panel's default correlator and with P = 32 (the
    table-driven correlator), against the oracle pixel by pixel."""
    import oracle
    W, H, T, cf = 7, 6, 8192, (1, 0)
    stack = imfcs.simulate_2d(W, H, T, 60, tau_bleach_s=15.0 if poly_order else 0.0, seed=41 + poly_order)
    bg = int(stack.min())
    a = _params(W, H, T, 1, cf, poly_order, bg=bg, p=p, q=q)
    pixels = plugin_glue.get_pixels(stack, a)
    bc = None
    if poly_order:
        bc, _ = plugin_glue.bleach_correction(pixels, a)
    lag, samp, _, _ = imfcs.correlator_tables(p, q, T, 1e-3)
    assert a.chanum == len(lag)
    n = a.width * a.height * a.chanum
    env = FakeEnv()
    outs = [np.zeros(n) for _ in range(3)]
    nb = [np.zeros(a.width * a.height) for _ in range(2)]
    bch = bc if bc is not None else np.zeros(1)
    env.call("calcACF", env.handle(pixels), env.handle(outs[0]), env.handle(outs[1]), env.handle(nb[0]),
             env.handle(nb[1]), env.handle(outs[2]), env.handle(bch), env.handle(samp),
             env.handle(lag.astype(np.int32)), _java_params(env, a))
    assert not env.exceptions and not env.trapped
    acf = outs[0].reshape(a.chanum, a.height, a.width)
    var = outs[1].reshape(a.chanum, a.height, a.width)
    ind = outs[2].reshape(a.chanum, a.height, a.width)[0]
    nc = a.bleachcorr_order
    worst = 0.0
    for y in range(a.height):
        for x in range(a.width):
            ca = cb = None
            if bc is not None:
                ca = bc[(y * a.w_temp + x) * nc:(y * a.w_temp + x + 1) * nc]
                cb = bc[(y * a.w_temp + x + cf[0]) * nc:(y * a.w_temp + x + cf[0] + 1) * nc]
            da = _oracle_series(stack, a, x, y, bg, ca)
            db = _oracle_series(stack, a, x + cf[0], y, bg, cb)
            oind, _ = oracle.block_transform(da, db, p=p, q=q)
            assert ind[y, x] == oind
            ocor, ovar, _, _ = oracle.calculate_cf(da, db, oind, p=p, q=q)
            amp = np.max(np.abs(ocor[1:]))
            e = rel_err(acf[:, y, x], ocor, 1e-4 * amp)
            worst = max(worst, e.max())
            np.testing.assert_allclose(var[1:, y, x], ovar[1:], rtol=1e-4)
    assert worst < 1e-4, worst


def test_calc_acf_nb_through_jni_matches_g1_statistics():
    """N&B (isNBcalculation): mean <n(t)> and lag-1 covariance
    <n(t) n(t+1)> - <n(t)><n(t+1)> of every pixel over frames 1..T-1, the
    statistics of the plugin's G1 N&B mode (IFCS:7872-7888), bleach
    corrected as in calcACF."""
    W, H, T = 6, 5, 8192
    stack = imfcs.simulate_2d(W, H, T, 50, tau_bleach_s=15.0, seed=61)
    bg = int(stack.min())
    a = _params(W, H, T, 1, (0, 0), 4, bg=bg)
    a.isNBcalculation = 1
    pixels = plugin_glue.get_pixels(stack, a)
    bc, _ = plugin_glue.bleach_correction(pixels, a)
    lag, samp, _, _ = imfcs.correlator_tables(16, 8, T, 1e-3)
    n = a.width * a.height * a.chanum
    env = FakeEnv()
    outs = [np.zeros(n) for _ in range(3)]
    nbm, nbc = np.zeros(a.width * a.height), np.zeros(a.width * a.height)
    env.call("calcACF", env.handle(pixels), env.handle(outs[0]), env.handle(outs[1]), env.handle(nbm),
             env.handle(nbc), env.handle(outs[2]), env.handle(bc), env.handle(samp), env.handle(lag.astype(np.int32)),
             _java_params(env, a))
    assert not env.exceptions and not env.trapped
    nc = a.bleachcorr_order
    for y in range(a.height):
        for x in range(a.width):
            k = y * a.w_temp + x
            d = _oracle_series(stack, a, x, y, bg, bc[k * nc:(k + 1) * nc])[1:]
            m1, m2 = d[:-1].mean(), d[1:].mean()
            cov = (d[:-1] * d[1:]).mean() - m1 * m2
            assert abs(nbm[y * a.width + x] - m1) <= 1e-9 * abs(m1)
            assert abs(nbc[y * a.width + x] - cov) <= 1e-4 * abs(cov)


def test_plugin_glue_acf_all_matches_native_pipeline(engine):
    """GPU_Calculate_ACF_All through the drop-in layer and the native
    pipeline give the same CCF for the same stack (no bleach correction:
    both paths then compute from identical series)."""
    W, H, T, cf = 8, 6, 8192, (1, 0)
    stack = imfcs.simulate_2d(W, H, T, 60, tau_bleach_s=0.0, seed=77)
    bg = int(stack.min())
    geom = imfcs.make_fit_model(PIX, PIX, PSF, rx=PIX)
    res = plugin_glue.calculate_acf_all(stack, W - 1, H, 1, cf, 1e-3, bg, 0, geom, [1.0, 1e-12, 0, 0, 0] + [0] * 6,
                                        [1, 1, 0, 0, 1] + [0] * 6, do_fit=False)
    cfg = imfcs.make_pipeline_config(W, H, T, binning=1, cf=cf, background=bg, poly_order=0, do_acf=False,
                                     do_fit=False)
    engine.upload(stack)
    engine.run(cfg)
    g = engine.results(cfg)
    ccf = np.moveaxis(res["acf"], 0, -1)
    amp = np.nanmax(np.abs(ccf[..., 1:]), axis=-1, keepdims=True)
    assert rel_err(g["acf"][0], ccf, 1e-4 * amp).max() < 1e-4
    np.testing.assert_array_equal(g["ind"][0], res["blocked"][0].astype(np.int32))


def test_fit_acf_1d_on_noisy_correlations_matches_java_fit():
    """The drop-in `fit` (ACF_1D, LSE, weights 1/variance, tol 1e-16f, 2000
    iterations: GPU_ACF_Fit, IFCS:15076-15256) on the correlations and
    blocked variances of simulated pixels, against the Java CPU fit
    (standardFit, IFCS:10288-10471) of the same curves.  N and D within 1e-3
    relative, G_inf within 1e-3 of 1/N, for fits CONVERGED on the GPU (the
    plugin keeps only those, IFCS:15249-15256)."""
    import oracle
    W = H = 8
    T = 20000
    stack = imfcs.simulate_2d(W, H, T, 60, tau_bleach_s=0.0, seed=303)
    bg = int(stack.min())
    om = oracle.make_model(PIX, PIX, PSF)
    o = oracle.pipeline(stack, oracle.make_config(W, H, T, background=bg, poly_order=0, do_fit=False, model=om,
                                                  nthreads=8))
    acf, var = o["acf"].reshape(-1, 73), o["var"].reshape(-1, 73)
    _, _, _, lagtime = imfcs.correlator_tables(16, 8, T, 1e-3)
    nfits, npts = acf.shape[0], 72
    geom = imfcs.make_fit_model(PIX, PIX, PSF)
    m = imfcs.FitModel(nfits, npts, True, imfcs.MODEL_ACF_1D, tolerance=np.float32(1e-16),
                       max_number_iterations=2000, num_valid_coefs=0,
                       parameters_to_fit=[True, True, False, False, True] + [False] * 15, user_info_size=npts * 4)
    m.data[:] = acf[:, 1:].reshape(-1).astype(np.float32)
    m.weights[:] = (np.float32(1.0) / var[:, 1:].astype(np.float32)).reshape(-1)
    row = [1.0, 1e-12, 0, 0, 0, 0, 0, 0, 0, 0, 0, geom.ax, geom.ay, geom.s, geom.sz, geom.rx, geom.ry,
           geom.fitobsvol, geom.q2, geom.q3]
    m.initial_parameters[:] = np.tile(np.asarray(row, np.float32), nfits)
    m.user_info[:npts] = lagtime[1:].astype(np.float32)
    r = imfcs.fit(m)
    p = r.parameters.reshape(nfits, 20)
    compared = 0
    for k in range(nfits):
        fr, chi2, ok, _, _ = oracle.standard_fit(om, lagtime, acf[k], var[k])
        if not ok or r.states[k] != 0:
            continue
        compared += 1
        assert abs(p[k, 0] - fr[0]) / fr[0] < 1e-3, (k, p[k, 0], fr[0])
        assert abs(p[k, 1] - fr[1]) / fr[1] < 1e-3, (k, p[k, 1], fr[1])
        assert abs(p[k, 4] - fr[4]) * fr[0] < 1e-3, (k, p[k, 4], fr[4])
    assert compared >= 0.9 * nfits, compared
