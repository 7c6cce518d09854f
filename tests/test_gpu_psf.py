"""The fit stage alone on stored correlations (imfcs_refit) and the PSF
calibration built on it (determinePSF, IFCS:8118-8160).

A refit with the run's own model reproduces the run's fits bit for bit; a
refit with another PSF size equals a fresh run with that size; the PSF
calibration's per-binning statistics equal the diffusion law computed
separately for each PSF size (the reference's procedure: dlfit per PSF).
"""
import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
from imaging_fcs_1_52_amd.difflaw import diffusion_law, psf_calibration
from parity import PIX

pytestmark = pytest.mark.gpu

FIT_KEYS = ("fitres", "chi2", "states", "iters")


def _cfg(W, H, T, bg, psf, b=1):
    return imfcs.make_pipeline_config(W, H, T, binning=b, background=bg, poly_order=4, do_fit=True,
                                      acf_model=imfcs.make_fit_model(b * PIX, b * PIX, psf))


def test_refit_matches_fresh_runs(engine):
    W, H, T = 24, 20, 6000
    stack = imfcs.simulate_2d(W, H, T, 120, D_um2s=1.0, tau_bleach_s=20.0, seed=77)
    bg = int(stack.min())
    d = engine.upload(stack)
    ca, cb = _cfg(W, H, T, bg, 0.38e-6), _cfg(W, H, T, bg, 0.30e-6)
    engine.run(ca, d)
    ra = engine.results(ca)
    engine.refit(ca)
    ra2 = engine.results(ca)
    for k in FIT_KEYS + ("acf", "var"):
        np.testing.assert_array_equal(np.asarray(ra2[k]), np.asarray(ra[k]), err_msg=k)
    engine.refit(cb)
    rb = engine.results(cb)
    fresh = imfcs.Engine(0)
    try:
        fresh.run(cb, d)
        rf = fresh.results(cb)
    finally:
        fresh.close()
    for k in FIT_KEYS + ("acf", "var"):
        np.testing.assert_array_equal(np.asarray(rb[k]), np.asarray(rf[k]), err_msg=k)
    assert not np.array_equal(np.asarray(rb["fitres"]), np.asarray(ra["fitres"]))
    # a configuration that changes the correlations is refused
    bad = _cfg(W, H, T, bg + 1, 0.38e-6)
    with pytest.raises(imfcs.ImFCSError):
        engine.refit(bad)


def test_psf_calibration_equals_diffusion_law_per_psf():
    W, H, T = 24, 24, 5000
    stack = imfcs.simulate_2d(W, H, T, 150, D_um2s=1.0, tau_bleach_s=20.0, seed=78)
    psfs = [0.34e-6, 0.40e-6]
    cal = psf_calibration(stack, psfs, 1, 3, PIX)
    assert cal["d_mean"].shape == (2, 3)
    np.testing.assert_array_equal(cal["label"], [1, 2, 3])
    for x, psf in enumerate(psfs):
        dl = diffusion_law(stack, 3, PIX, psf)
        np.testing.assert_array_equal(cal["d_mean"][x], dl["d_mean"])
        np.testing.assert_array_equal(cal["d_var"][x], dl["d_var"])
        np.testing.assert_array_equal(cal["n_fitted"][x], dl["n_fitted"])
        for i, b in enumerate((1, 2, 3)):
            assert cal["err"][x, i] == pytest.approx(dl["d_var"][i] / np.sqrt(((W // b) * H) // b), rel=1e-12)
    # the trial PSF size changes the fitted D
    assert np.all(cal["d_mean"][1] != cal["d_mean"][0])
