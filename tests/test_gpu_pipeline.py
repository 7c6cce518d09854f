"""GPU parity of the native pipeline against the oracle (the Java CPU path
restated in C) on identical seeded stacks.

Tolerances (stated by the engine's contract): correlation channels within
1e-4 relative (floor 1e-4 of the curve amplitude for channels near zero),
blocked variances within 1e-4 relative where the blocking index agrees,
fitted N, D within 1e-3 relative and G_inf within 1e-3 of the amplitude
1/N, for pixels fitted on both sides.
"""
import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
from parity import rel_err, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _stack(W=8, H=8, T=8192, n=70, seed=11, tau=15.0, D=1.0):
    return imfcs.simulate_2d(W, H, T, n, D_um2s=D, tau_bleach_s=tau, seed=seed)


def _compare_corr(g_acf, g_var, g_ind, o):
    amp = np.nanmax(np.abs(o["acf"][..., 1:]), axis=-1, keepdims=True)
    finite = np.isfinite(o["acf"]).all(-1)
    assert np.array_equal(np.isfinite(g_acf).all(-1), finite), "NaN pattern differs"
    e = rel_err(g_acf[finite], o["acf"][finite], 1e-4 * amp[finite])
    assert e.max() < 1e-4, e.max()
    same = (g_ind == o["ind"]) & finite
    assert same[finite].mean() >= 0.97, same[finite].mean()
    ev = rel_err(g_var[same][:, 1:], o["var"][same][:, 1:], 1e-300)
    assert ev.max() < 1e-4, ev.max()
    return same


def _compare_fit(gres, o, same, pair):
    gfit = gres["fitres"][pair]
    ok = (gres["states"][pair] == 0) & o["fitted"] & same
    assert ok.mean() > 0.5, ok.mean()
    N_g, N_o = gfit[..., 0][ok], o["fitres"][..., 0][ok]
    D_g, D_o = gfit[..., 1][ok], o["fitres"][..., 1][ok]
    G_g, G_o = gfit[..., 4][ok], o["fitres"][..., 4][ok]
    assert rel_err(N_g, N_o, 1e-12).max() < 1e-3
    assert rel_err(D_g, D_o, 1e-30).max() < 1e-3
    assert (np.abs(G_g - G_o) / (1.0 / np.abs(N_o))).max() < 1e-3


@pytest.mark.parametrize("binning", [1, 2])
def test_acf_ccf_bleach_poly4(engine, binning):
    stack = _stack(W=10, H=8, seed=5 + binning)
    cf = (binning, 0)
    gres, cfg = run_gpu(engine, stack, binning=binning, cf=cf, poly_order=4)
    o_acf = run_oracle(stack, binning=binning, cf=(0, 0), poly_order=4, roi_cf=cf)
    o_ccf = run_oracle(stack, binning=binning, cf=cf, poly_order=4)
    same = _compare_corr(gres["acf"][0], gres["var"][0], gres["ind"][0], o_acf)
    _compare_fit(gres, o_acf, same, 0)
    same_c = _compare_corr(gres["acf"][1], gres["var"][1], gres["ind"][1], o_ccf)
    assert same_c.any()


def test_acf_no_bleach_correction(engine):
    stack = _stack(seed=21, tau=0.0)
    gres, _ = run_gpu(engine, stack, poly_order=0)
    o = run_oracle(stack, poly_order=0)
    same = _compare_corr(gres["acf"][0], gres["var"][0], gres["ind"][0], o)
    _compare_fit(gres, o, same, 0)


def test_polynomial_coefficients_match_least_squares(engine):
    """The bleach polynomial is the least-squares optimum of the integer
    trace; the oracle's LM (PolynomFit) converges to the same optimum."""
    import oracle
    stack = _stack(seed=31)
    gres, _ = run_gpu(engine, stack, poly_order=4, do_fit=False)
    T = stack.shape[0]
    bg = int(stack.min())
    for (y, x) in [(0, 0), (3, 5), (7, 7)]:
        tr, it = oracle.intensity_trace(stack, x, y, 1, 1, bg, 1, T, 1000, 1e-3)
        ref = np.polyfit(it, tr, 4)[::-1]
        k = y * 8 + x
        np.testing.assert_allclose(gres["coef"][k], ref, rtol=1e-7, atol=1e-9 * abs(ref[0]))


def test_ccf_fits_take_the_ccf_models_free_parameters(engine):
    """ACF fits take acf_model.paramfit and CCF fits ccf_model.paramfit
    (GPU_ACF_Fit fits every curve with the panel's hold flags,
    IFCS:15131-15141): with different free sets the ACF and the CCF results
    equal those of runs where both models carry that kind's set."""
    stack = _stack(W=8, H=6, seed=91)
    T, H, W = stack.shape
    bg = int(stack.min())
    pa = [1, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0]
    pc = [1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0]   # D held at its initial value for the CCF
    ini = [1.0, 0.8e-12, 0, 0, 0, 0, 0, 0, 0, 0, 0]

    def run(fa, fc):
        acf = imfcs.make_fit_model(0.24e-6, 0.24e-6, 0.38e-6, paramfit=fa, paraminit=ini)
        ccf = imfcs.make_fit_model(0.24e-6, 0.24e-6, 0.38e-6, rx=0.24e-6, paramfit=fc, paraminit=ini)
        cfg = imfcs.make_pipeline_config(W, H, T, cf=(1, 0), background=bg, poly_order=4, acf_model=acf,
                                         ccf_model=ccf)
        engine.upload(stack)
        engine.run(cfg)
        return engine.results(cfg)

    mixed, both_a, both_c = run(pa, pc), run(pa, pa), run(pc, pc)
    assert np.array_equal(mixed["fitres"][0], both_a["fitres"][0], equal_nan=True)
    assert np.array_equal(mixed["fitres"][1], both_c["fitres"][1], equal_nan=True)
    assert np.array_equal(mixed["chi2"][1], both_c["chi2"][1], equal_nan=True)
    assert np.array_equal(mixed["states"], np.stack([both_a["states"][0], both_c["states"][1]]))
    ok = mixed["states"][1] == 0
    assert ok.any()
    np.testing.assert_allclose(mixed["fitres"][1][..., 1][ok], 0.8e-12, rtol=1e-6)


def test_pinned_host_buffer_round_trip(engine):
    """imfcs_host_alloc: a stack staged in pinned host memory reaches the
    device unchanged, and a run on it equals a run on the pageable upload."""
    stack = imfcs.simulate_2d(10, 8, 4096, 30, seed=7)
    T, H, W = stack.shape
    pin = engine.pinned_array(stack.shape)
    pin[...] = stack
    d = engine.upload(stack)
    cfg = imfcs.make_pipeline_config(W, H, T, background=int(stack.min()), poly_order=4, do_fit=False)
    engine.run(cfg, d)
    ref = engine.results(cfg)["acf"].copy()
    d_pin = engine.device_buffer(stack.nbytes)
    engine.copy_in(d_pin, pin)
    assert np.array_equal(engine.download(d_pin, stack.shape), stack)
    engine.run(cfg, d_pin)
    assert np.array_equal(engine.results(cfg)["acf"], ref, equal_nan=True)
    engine.free_buffer(d_pin)
    engine.free_pinned(pin)
