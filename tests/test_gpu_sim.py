"""The GPU particle simulator (imfcs_simulate_2d_device).

Checks what its physics fixes in closed form:
  * a static emitter (D = 0, no bleaching) gives every pixel Poisson counts
    with mean lambda * (PSF mass over the pixel): mean == variance per pixel,
    and the means summed over the 7 x 7 footprint equal lambda (mass inside
    > 0.999), on both Poisson paths (lambda = 5: multiplication method;
    lambda = 1500: transformed rejection);
  * a homogeneous periodic field has mean counts per pixel and frame of
    density x brightness, and a brightness ramp follows counts_lo ->
    counts_hi along x;
  * bleaching removes emitters at rate 1/tau (mean intensity e^{-t/tau});
  * the same seed gives the same stack, another seed another one;
  * diffusion: the plugin's model fitted to the pipeline's ACFs returns D.
"""
import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs

pytestmark = pytest.mark.gpu

PIX = 0.24


@pytest.mark.parametrize("lam", [5.0, 1500.0])
def test_static_emitter_counts_are_poisson_with_psf_means(engine, lam):
    W = H = 16
    T = 20000
    # one emitter inside the central 2 x 2 pixels: its 7 x 7 footprint is on the detector
    dom = imfcs.sim_domain(7 * PIX, 7 * PIX, 2 * PIX, 2 * PIX, 1, D_um2s=0.0, counts=lam)
    d = imfcs.simulate_2d_device(engine, W, H, T, dom, offset=0, tau_bleach_s=0.0, seed=5)
    st = engine.download(d, (T, H, W)).astype(np.float64)
    m, v = st.mean(0), st.var(0)
    assert abs(m.sum() - lam) < 4 * np.sqrt(lam / T) + 1e-3 * lam, (m.sum(), lam)
    big = m > 0.05 * m.max()
    # per-pixel mean == variance within the sampling error of the variance
    rel = np.abs(v[big] - m[big]) / m[big]
    assert rel.max() < 6 * np.sqrt(2.0 / T) + 0.02, rel.max()
    # the footprint is a separable Gaussian: the marginals have sigma 0.19 / 0.24 pixel
    mx = m.sum(0) / m.sum()
    xs = np.arange(W) + 0.5
    mu = (mx * xs).sum()
    sd = np.sqrt((mx * (xs - mu) ** 2).sum())
    # pixelation adds 1/12 pixel^2 of variance
    assert abs(sd ** 2 - ((0.19 / PIX) ** 2 + 1.0 / 12)) < 0.05


def test_homogeneous_field_mean_ramp_bleaching_and_seed(engine):
    W, H, T = 64, 48, 4000
    m = 2
    n = 3000
    dom = imfcs.sim_domain(-m * PIX, -m * PIX, (W + 2 * m) * PIX, (H + 2 * m) * PIX, n, D_um2s=1.0, counts=20.0)
    d = imfcs.simulate_2d_device(engine, W, H, T, dom, offset=100, tau_bleach_s=0.0, seed=11)
    a = engine.download(d, (T, H, W))
    dens = n / ((W + 2 * m) * (H + 2 * m))
    mean = a.astype(np.float64).mean() - 100
    assert abs(mean / (dens * 20.0) - 1) < 0.03, (mean, dens * 20)
    d = imfcs.simulate_2d_device(engine, W, H, T, dom, offset=100, tau_bleach_s=0.0, seed=11)
    assert np.array_equal(engine.download(d, (T, H, W)), a)
    d = imfcs.simulate_2d_device(engine, W, H, T, dom, offset=100, tau_bleach_s=0.0, seed=12)
    assert not np.array_equal(engine.download(d, (T, H, W)), a)
    # bleaching: mean intensity e^{-t/tau}
    d = imfcs.simulate_2d_device(engine, W, H, T, dom, offset=100, tau_bleach_s=2.0, seed=13)
    b = engine.download(d, (T, H, W)).astype(np.float64).mean((1, 2)) - 100
    t = (np.arange(T) + 1) * 1e-3
    ratio = b[-500:].mean() / b[:500].mean()
    expect = np.exp(-t[-500:] / 2.0).mean() / np.exp(-t[:500] / 2.0).mean()
    assert abs(ratio / expect - 1) < 0.05, (ratio, expect)
    # brightness ramp 10 -> 2000 along x (slowly diffusing emitters)
    dom = imfcs.sim_domain(-m * PIX, -m * PIX, (W + 2 * m) * PIX, (H + 2 * m) * PIX, n, D_um2s=0.01, counts=10.0,
                           counts_hi=2000.0)
    d = imfcs.simulate_2d_device(engine, W, H, 2000, dom, offset=100, tau_bleach_s=0.0, seed=14)
    c = engine.download(d, (2000, H, W)).astype(np.float64).mean((0, 1)) - 100
    left, right = c[4:8].mean(), c[-8:-4].mean()
    assert 50 < right / left < 400, (left, right)


def test_simulated_diffusion_is_recovered_by_the_pipeline(engine):
    W = H = 24
    T = 20000
    m = 2
    D = 2.0
    dom = imfcs.sim_domain(-m * PIX, -m * PIX, (W + 2 * m) * PIX, (H + 2 * m) * PIX, 400, D_um2s=D, counts=30.0)
    d = imfcs.simulate_2d_device(engine, W, H, T, dom, offset=100, tau_bleach_s=0.0, seed=21)
    acf = imfcs.make_fit_model(PIX * 1e-6, PIX * 1e-6, 0.38e-6)
    cfg = imfcs.make_pipeline_config(W, H, T, background=100, poly_order=0, acf_model=acf)
    engine.run(cfg, d)
    r = engine.results(cfg)
    ok = r["states"][0] == 0
    assert ok.mean() > 0.9
    Dfit = np.median(r["fitres"][0][..., 1][ok]) * 1e12
    assert abs(Dfit / D - 1) < 0.15, Dfit
